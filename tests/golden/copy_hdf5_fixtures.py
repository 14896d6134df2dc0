"""Copy a selection of the reference's HDF5 regression data into tests/golden/hdf5.

Run once in the development container (``python tests/golden/copy_hdf5_fixtures.py``).  Only
data files are copied -- the fnetdata datasets, the input netstats (``fortnet.hdf5``), the
expected netstats after training (``_fortnet.hdf5``) and the expected outputs
(``_fnetout.hdf5``) that the reference's own tests hold under ``test/prog/fortnet`` -- so the
HDF5 layer can be checked against files the reference wrote, on any machine.  A case directory
``a/b/c`` becomes ``hdf5/a__b__c``, the naming of the converted ``.npz`` cases.
"""
import os
import shutil

REF = "/root/reference/test/prog/fortnet"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hdf5")

CASES = [
    "analysis/derivatives/analytical/%s/%s/%s" % (t, s, k)
    for t in ("singleTarget", "multiTargets")
    for s in ("speciesResolved", "speciesUnresolved")
    for k in ("raw", "atomID")
] + [
    "prediction/multispecies/globalTargets/acsffeatures/speciesResolved",
    "validation/multispecies/bothTargets",
] + [
    "restart/%s/%s/optimizer/steepestdesc" % (s, t)
    for s in ("singlespecies", "multispecies")
    for t in ("globalTargets", "atomicTargets")
]
FILES = ("fortnet.hdf5", "_fortnet.hdf5", "_fnetout.hdf5")


def main():
    os.makedirs(os.path.join(OUT, "datasets"), exist_ok=True)
    for fn in sorted(os.listdir(os.path.join(REF, "datasets"))):
        shutil.copyfile(os.path.join(REF, "datasets", fn), os.path.join(OUT, "datasets", fn))
    for case in CASES:
        src = os.path.join(REF, case)
        dst = os.path.join(OUT, case.replace("/", "__"))
        os.makedirs(dst, exist_ok=True)
        for fn in FILES:
            if not os.path.exists(os.path.join(src, fn)):
                continue
            # training runs rewrite fortnet.hdf5; prediction runs only read it
            if fn == "_fortnet.hdf5" and not case.startswith("restart"):
                continue
            shutil.copyfile(os.path.join(src, fn), os.path.join(dst, fn))


if __name__ == "__main__":
    main()
