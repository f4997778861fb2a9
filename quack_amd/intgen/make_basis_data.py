"""Extract the basis-set data the producer needs into basis_data.json.

Reads the NWChem-format files that QuAcK ships (basis/cc-pvdz, basis/6-31g;
the same files PyDuck.py:108-113 hands to PySCF) and keeps the elements used
by the workloads here (H, He, O).  Run once in a tree that holds the QuAcK
sources; the JSON it writes is committed so that nothing reads QuAcK's tree at
run time.

    python quack_amd/intgen/make_basis_data.py /root/reference/basis
"""
import json
import os
import sys

L_OF = {"S": 0, "P": 1, "D": 2}


def parse_nwchem(path, elements):
    out = {el: [] for el in elements}
    cur = None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#") or t[0] in ("BASIS", "END"):
                if t and t[0] == "END":
                    cur = None
                continue
            if t[0] in elements and len(t) == 2 and t[1] in L_OF:
                cur = {"l": L_OF[t[1]], "exps": [], "coefs": []}
                out[t[0]].append(cur)
            elif t[0][0].isalpha():
                cur = None
            elif cur is not None:
                cur["exps"].append(float(t[0].replace("D", "E")))
                cur["coefs"].append(float(t[1].replace("D", "E")))
    return out


def main(basis_dir):
    data = {}
    for name in ("cc-pvdz", "6-31g"):
        data[name] = parse_nwchem(os.path.join(basis_dir, name), ["H", "He", "O"])
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "basis_data.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/basis")
