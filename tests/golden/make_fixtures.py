"""Extract the reference's literal D3Q19 tables as data fixtures.

Reads main.cpp and kernel.cu from the reference checkout as text (nothing is
compiled or run) and writes:
  mrt_matrix_main_cpp.txt   the 19x19 moment matrix literal of main.cpp:150-171
  mrt_matrix_kernel_cu.txt  the same literal of kernel.cu:63-84
  d3q19_tables.txt          Direction[XDIM|YDIM|ZDIM][q] (main.cpp:87-105),
                            OppIndex (main.cpp:107-113)
  common_h_constants.txt    the case numbers and node flags of common.h:26-37
Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def read(ref, name):
    with open(os.path.join(ref, name), "rb") as f:
        txt = f.read().decode("utf-8-sig")
    return txt.replace("\r\n", "\n")


def matrix_literal(txt, start_pat):
    i = txt.index(start_pat)
    j = txt.index("};", i)
    rows = re.findall(r"\{([^{}]*)\}", txt[i:j + 2])
    m = [[float(v) for v in r.split(",")] for r in rows]
    assert len(m) == 19 and all(len(r) == 19 for r in m), "unexpected matrix literal"
    return m


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    main_cpp = read(ref, "main.cpp")
    kernel_cu = read(ref, "kernel.cu")
    for name, m in (("mrt_matrix_main_cpp.txt", matrix_literal(main_cpp, "mat=")),
                    ("mrt_matrix_kernel_cu.txt", matrix_literal(kernel_cu, "mat =\n"))):
        with open(os.path.join(HERE, name), "w") as f:
            f.write("# 19x19 moment matrix literal, row k = moment k, column q = direction q\n")
            for r in m:
                f.write(" ".join(f"{v:g}" for v in r) + "\n")
    d = {0: {}, 1: {}, 2: {}}
    axis = {"XDIM": 0, "ZDIM": 1, "YDIM": 2}
    for a, q, v in re.findall(r"m_HostDirection\[(XDIM|YDIM|ZDIM)\]\[(\d+)\]\s*=\s*(-?[0-9.]+)f", main_cpp):
        d[axis[a]][int(q)] = float(v)
    opp = {int(q): int(v) for q, v in re.findall(r"m_HostOppIndex\[(\d+)\]\s*=\s*(\d+);", main_cpp)}
    assert all(len(d[a]) == 19 for a in d) and len(opp) == 19
    with open(os.path.join(HERE, "d3q19_tables.txt"), "w") as f:
        f.write("# rows: Direction[XDIM=0][q], Direction[ZDIM=1][q], Direction[YDIM=2][q], OppIndex[q]\n")
        for a in (0, 1, 2):
            f.write(" ".join(f"{d[a][q]:g}" for q in range(19)) + "\n")
        f.write(" ".join(str(opp[q]) for q in range(19)) + "\n")
    common = read(ref, "common.h")
    consts = {}
    for name in ("CAVITY", "POISEUILLE", "NODE_FLUID", "NODE_FIXEDWALL", "NODE_MOVINGWALL", "NODE_INLET",
                 "NODE_OUTLET", "NODE_STRUCT"):
        m = re.search(r"#define\s+" + name + r"\s+(?:\(1 << (\d+)\)|(\d+))", common)
        assert m, name
        consts[name] = 1 << int(m.group(1)) if m.group(1) else int(m.group(2))
    with open(os.path.join(HERE, "common_h_constants.txt"), "w") as f:
        f.write("# name value, from the #defines of common.h\n")
        for k, v in consts.items():
            f.write(f"{k} {v}\n")
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
