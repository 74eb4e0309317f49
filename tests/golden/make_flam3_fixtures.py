"""Copy a subset of the reference's flam3-C parity corpus into tests/golden/flam3.

Each fixture directory holds the source `<id>.flam3`, the flam3-C render
`golden.png` and `meta.json` (size, expectedR / thresholdR: the reference's
own mean |difference| over RGBA bytes and its pass bar), unchanged. The
subset covers: final xforms, dihedral symmetry, supersample 4, DE radius 0,
a 1280x720 frame, a two-keyframe file, tier-1 and tier-2 fixtures.
Attribution: Electric Sheep genomes (CC-BY / CC-BY-NC 2.0; see README.md
in each fixture set of the reference, fixtures/flam3-goldens/README.md).

    python tests/golden/make_flam3_fixtures.py /root/reference
"""
import shutil
import sys
from pathlib import Path

SUBSET = ["244.57686", "coverage.245.06687", "244.00016", "244.00617", "coverage.247.20817",
          "coverage.247.31007", "248.11268", "248.23554", "248.25703", "247.29388"]


def main(ref: str) -> None:
    src = Path(ref) / "fixtures" / "flam3-goldens"
    dst = Path(__file__).resolve().parent / "flam3"
    dst.mkdir(exist_ok=True)
    for fid in SUBSET:
        d = dst / fid
        d.mkdir(exist_ok=True)
        for name in (f"{fid}.flam3", "golden.png", "meta.json"):
            shutil.copyfile(src / fid / name, d / name)
    print(f"copied {len(SUBSET)} fixtures into {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
