"""Regenerate ``new-joe-kuo-6.21201.4096.txt`` (the published Joe--Kuo Sobol
direction-number dataset, first 4096 table rows = Sobol dimensions 2..4097).

The dataset is S. Joe and F. Y. Kuo, "new-joe-kuo-6.21201"
(https://web.maths.unsw.edu.au/~fkuo/sobol/).  It is written here in the
dataset's own published text layout, one dimension per line::

    d  s  a  m_1 .. m_s

The values are read, as numbers only, from the table the reference engine
embeds (pricer/src/nodes/sobol_direction_numbers.cpp:16-70004, arrays S, A,
M_OFFSET, M), so that this engine uses exactly the same direction numbers.
Run once by hand; the output file is committed data.
"""
import re
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else
           "/root/reference/pricer/src/nodes/sobol_direction_numbers.cpp")
OUT = Path(__file__).with_name("new-joe-kuo-6.21201.4096.txt")


def array(text, name):
    m = re.search(r"sobol_jk::%s\[\d+\]\s*=\s*\{(.*?)\};" % name, text, re.S)
    return [int(v) for v in re.findall(r"\d+", m.group(1))]


def main():
    text = SRC.read_text()
    s, a, off, m = (array(text, k) for k in ("S", "A", "M_OFFSET", "M"))
    assert len(s) == len(a) == 4096 and len(off) == 4097 and off[-1] == len(m)
    lines = ["d s a m_i"]
    for i in range(4096):
        mi = m[off[i]:off[i + 1]]
        assert len(mi) == s[i]
        lines.append(" ".join(str(v) for v in [i + 2, s[i], a[i]] + mi))
    OUT.write_text("\n".join(lines) + "\n")
    print("wrote", OUT, len(lines) - 1, "dimensions")


if __name__ == "__main__":
    main()
