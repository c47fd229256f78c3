"""Writes the golden fixtures of tests/golden/ from data files of the reference tree.

    python tests/golden/make_golden.py [/root/reference]

stencils_ref.npz: the reference's element and restriction stencil tables, parsed from
  Stencils/LapType2Stencils.inp, MassType2Stencils.inp, RmatType2Stencils.inp and
  RmatType1Stencils.inp with the same whitespace-separated "%lf" reading as
  examples/src/backend/handleType2Stencils.C:776-794 and src/omg/RmatType{1,2}StencilsAux.C.
stencils_ref2.npz: LapType1Stencils.inp, MassType1Stencils.inp ([coarse child][eType][fine
  child][8][8], examples/src/backend/handleType1Stencils.C:439-520), ShFnStencils.inp
  ([child][eType][vertex][3][3][3], handleType2Stencils.C:212-225) and GDType2Stencils.inp
  ([child][eType][24][24], handleType2Stencils.C:315-323), read the same way.
etype_masks.json: the 18 hanging masks of each child number in eType order, parsed from
  the enum values of include/oda/cnumEtypes.h.
Only data values are extracted (numbers), never source text.
"""
import json
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def table(name, shape):
    vals = np.array(open(os.path.join(REF, "Stencils", name)).read().split(), dtype=np.float64)
    assert vals.size == int(np.prod(shape)), (name, vals.size)
    return vals.reshape(shape)


def main():
    np.savez_compressed(
        os.path.join(OUT, "stencils_ref.npz"),
        lap2=table("LapType2Stencils.inp", (8, 18, 8, 8)),
        mass2=table("MassType2Stencils.inp", (8, 18, 8, 8)),
        rmat2=table("RmatType2Stencils.inp", (8, 18, 8, 8)),
        rmat1=table("RmatType1Stencils.inp", (8, 8, 18, 8, 8)),
    )
    np.savez_compressed(
        os.path.join(OUT, "stencils_ref2.npz"),
        lap1=table("LapType1Stencils.inp", (8, 18, 8, 8, 8)),
        mass1=table("MassType1Stencils.inp", (8, 18, 8, 8, 8)),
        shfn=table("ShFnStencils.inp", (8, 18, 8, 3, 3, 3)),
        gd2=table("GDType2Stencils.inp", (8, 18, 24, 24)),
    )
    txt = open(os.path.join(REF, "include", "oda", "cnumEtypes.h")).read()
    parts = re.split(r"class cNumEtype<(\d)>", txt)
    masks = {}
    for i in range(1, len(parts), 2):
        masks[int(parts[i])] = [int(v) for v in re.findall(r"ET_\w+\s*=\s*(\d+)", parts[i + 1])][:18]
    assert sorted(masks) == list(range(8)) and all(len(v) == 18 for v in masks.values())
    with open(os.path.join(OUT, "etype_masks.json"), "w") as f:
        json.dump({str(k): masks[k] for k in range(8)}, f, indent=1)


if __name__ == "__main__":
    main()
