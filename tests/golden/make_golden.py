"""Extract golden vectors from the reference's validation logs into JSON.

Run in a container that has the reference checkout (default /root/reference):
    python tests/golden/make_golden.py [/path/to/reference]
Writes tests/golden/validation_logs.json.  Only numbers are copied, each with
the log file and line range it comes from.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "validation_logs.json")


def rows(path, first, last):
    """numeric rows of lines first..last (1-based, inclusive)"""
    with open(os.path.join(REF, path)) as f:
        lines = f.read().split("\n")
    out = []
    for ln in lines[first - 1:last]:
        vals = []
        for tok in ln.replace("nod", " ").split():
            try:
                vals.append(float(tok.replace("D", "E")))
            except ValueError:
                pass
        out.append(vals)
    return out


def main():
    one = "validation/1elem_3d_red_int_f_0.06.txt"
    cube = "validation/1step_red_int_cube3D_hf_c_0.06.txt"
    g = {
        "ch_constants": {
            "source": f"{one}:112-114 (C++ printout)",
            "alpha_beta_gamma": [r[0] for r in rows(one, 112, 114)],
        },
        "cube_dHdetJ": {
            "source": f"{cube}:7-9 (F90, 0.1 m cube, rows d = x, y, z, columns local node 0..7)",
            "h": 0.1,
            "values": rows(cube, 7, 9),
        },
        "one_elem_hg": {
            "source": f"{one}:7-14 displacements, :16-23 velocities (C++, 7 digits)",
            "case": "0.1 m cube, E=206e9, nu=0.3, rho=7850, elastic, dt=0.8e-5, tf=1e-3, "
                    "node 0 fixed xyz, 1 fixed yz, 2 fixed xz, 3 fixed z, nodes 4-7 vz=-1, hourglass 0.06",
            "u": rows(one, 7, 14),
            "v": rows(one, 16, 23),
        },
        "one_elem_no_hg": {
            "source": f"{one}:76-83 displacements, :85-92 velocities (C++, 7 digits)",
            "u": rows(one, 76, 83),
            "v": rows(one, 85, 92),
        },
    }
    with open(OUT, "w") as f:
        json.dump(g, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
