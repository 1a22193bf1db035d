"""Copy the reference's own regression data for the NS/SA path into tests/golden/reference.

The files copied are data the reference's test harness (test/test.py) already
holds: the case grids (.ofl, binary), the golden outputs in each case's
autotest/ folder, and the restart file (restart/flow.dat, binary) the
continued plate case starts from.  No reference source file is copied.  Run from the
repository root in a container that has the reference tree:

    python tests/golden/collect_reference_fixtures.py /root/reference
"""
import shutil
import sys
from pathlib import Path

CASES = {
    # case dir: (grid file, golden files)
    "plateuns2dslau2": ("grid/plate20182d.ofl", ["res.dat", "aero.dat", "wallaero.dat"]),
    "turbplateuns2droe_sa": ("grid/turbplate20182d.ofl", ["res.dat", "turbres.dat", "aero.dat", "wallaero.dat"]),
    "rae2822_roe_sa": ("grid/rae2822_2d.ofl", ["res.dat", "turbres.dat", "aero.dat", "wallaero.dat"]),
    # restart from step 34950 of the plate (same grid as plateuns2dslau2), steps 34951..35000
    "plateuns2dslau2_continue": (None, ["res.dat", "aero.dat"]),
}
RESTARTS = {"plateuns2dslau2_continue": "restart/flow.dat"}


def main(ref_root: str) -> None:
    out = Path(__file__).resolve().parent / "reference"
    for case, (grid, goldens) in CASES.items():
        src = Path(ref_root) / "test" / case
        dst = out / case
        dst.mkdir(parents=True, exist_ok=True)
        if grid:
            shutil.copyfile(src / grid, dst / "grid.ofl")
        if case in RESTARTS:
            shutil.copyfile(src / RESTARTS[case], dst / "flow.dat")
        for g in goldens:
            shutil.copyfile(src / "autotest" / g, dst / g)
        print("copied", case)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
