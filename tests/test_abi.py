"""The device library loads (no GPU needed) and exports every C-ABI entry point
declared in include/*.h; without a device the gate reports 'unavailable'
instead of falling back to a CPU path."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(thoth_[a-z_0-9]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_the_boundary():
    names = declared_functions()
    assert {"thoth_gpu_available", "thoth_gpu_device_info", "thoth_gpu_price_european_gbm",
            "thoth_mcl_sweep", "thoth_mcl_price", "thoth_last_error", "thoth_gpu_sobol_points",
            "thoth_cv_train", "thoth_cv_price", "thoth_cv_loss_grad"} <= names


def test_library_exports_every_declared_symbol():
    from thoth_amd import mcl
    lib = mcl.lib()
    exported = subprocess.run(["nm", "-D", "--defined-only", str(mcl.LIB)], capture_output=True, text=True).stdout
    for name in declared_functions():
        assert getattr(lib, name) is not None
        assert re.search(r"\bT %s\b" % name, exported), name


def test_structs_match_header_sizes(tmp_path):
    # compile a probe of the header struct sizes and compare with the ctypes mirror
    from thoth_amd import abi
    structs = [("thoth_mcl_request", abi.Request), ("thoth_mcl_result", abi.Result), ("thoth_asset", abi.Asset),
               ("thoth_contract", abi.Contract), ("thoth_mcl_book", abi.Book), ("thoth_mcl_sweep_cfg", abi.SweepCfg),
               ("thoth_mcl_book_out", abi.BookOut), ("thoth_cv_weights", abi.CvWeights),
               ("thoth_cv_config", abi.CvConfig), ("thoth_cv_train_report", abi.CvTrainReport),
               ("thoth_cv_result", abi.CvResult), ("thoth_rqmc_result", abi.RqmcResult),
               ("thoth_pathwise_result", abi.PathwiseResult), ("thoth_vanilla_grid", abi.VanillaGrid),
               ("thoth_grid_result", abi.GridResult), ("thoth_grid_tuning", abi.GridTuning),
               ("thoth_plan_shape", abi.PlanShape)]
    src = tmp_path / "abi_probe.c"
    src.write_text('#include <stdio.h>\n#include "thoth_mcl_gpu.h"\n#include "thoth_mcl_cv.h"\n#include "thoth_mcl_grid.h"\nint main(){'
                   + "".join('printf("%%zu\\n", sizeof(%s));' % n for n, _ in structs) + "}\n")
    exe = tmp_path / "abi_probe"
    subprocess.run(["gcc", "-I%s" % (ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert got == [C.sizeof(t) for _, t in structs]


def test_no_device_means_unavailable_not_fallback():
    from thoth_amd import mcl
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    assert not mcl.available()
    from thoth_amd import configs
    from thoth_amd.book import BuiltRequest
    with pytest.raises(mcl.Unavailable):
        mcl.price(BuiltRequest(configs.black_scholes_call(paths=1024)))
