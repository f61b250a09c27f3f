"""
Extract the NIST Igor SANS comparison data that the reference's resolution
tests hold (sasmodels/resolution.py:1046-1122: Q, dQ, I(Q) for the pinhole
sphere; Q, dQv, I(Q), I_smeared(Q) for the slit sphere) into plain-text
fixtures, with the sphere parameters of each data set.

    python tests/golden/make_resolution_fixtures.py [/root/reference]

The reference file is read as text; nothing of it is imported or run.
"""

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def block(text, name):
    m = re.search(name + r' = """\\\n(.*?)"""', text, re.S)
    return m.group(1).strip() + "\n"


def pars(text, name):
    m = re.search(name + r" = \{(.*?)\}", text, re.S)
    return {k: float(v) for k, v in re.findall(r"'(\w+)':\s*([-\d.eE+]+)", m.group(1))}


def main(root="/root/reference"):
    text = open(os.path.join(root, "sasmodels", "resolution.py")).read()
    for kind in ("PINHOLE", "SLIT"):
        data = block(text, "TEST_DATA_%s_SPHERE" % kind)
        with open(os.path.join(HERE, "igor_%s_sphere.txt" % kind.lower()), "w") as f:
            f.write(data)
        with open(os.path.join(HERE, "igor_%s_sphere.json" % kind.lower()), "w") as f:
            json.dump(pars(text, "TEST_PARS_%s_SPHERE" % kind), f, indent=1)


if __name__ == "__main__":
    main(*sys.argv[1:])
