"""Copy the reference's flam3-C parity corpus into tests/golden/flam3.

Each fixture directory holds the source `<id>.flam3`, the flam3-C render
`golden.png` and `meta.json` (size, expectedR / thresholdR: the reference's
own mean |difference| over RGBA bytes and its pass bar), unchanged: all 26
fixtures the reference's gate runs (src/parity.test.ts:117-148), tier 1 and
tier 2 (final xforms, dihedral symmetry, supersample 4, DE radius 0,
1280x720 frames, two-keyframe files).
Attribution: Electric Sheep genomes (CC-BY / CC-BY-NC 2.0; see README.md
in each fixture set of the reference, fixtures/flam3-goldens/README.md).

    python tests/golden/make_flam3_fixtures.py /root/reference
"""
import shutil
import sys
from pathlib import Path


def main(ref: str) -> None:
    src = Path(ref) / "fixtures" / "flam3-goldens"
    dst = Path(__file__).resolve().parent / "flam3"
    dst.mkdir(exist_ok=True)
    ids = sorted(d.name for d in src.iterdir() if (d / "meta.json").exists())
    for fid in ids:
        d = dst / fid
        d.mkdir(exist_ok=True)
        for name in (f"{fid}.flam3", "golden.png", "meta.json"):
            shutil.copyfile(src / fid / name, d / name)
    print(f"copied {len(ids)} fixtures into {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
