"""Copies the reference's own scene (scenes/obj-model.txt) and the OBJ models its first 12 objects
use into tests/golden/reference_scene/, keeping the reference's relative layout (the scene names its
models as models/<file>).  These are data files of the reference repository, used unchanged as
inputs of the scene-ingest tests.  Run from the repository root while the reference tree is present:

    python tests/golden/make_reference_scene.py /root/reference
"""
import shutil
import sys
from pathlib import Path

MODELS = ["plane-x.obj", "plane-y.obj", "plane-z.obj", "cube.obj", "cube.mtl", "table.obj", "teapot.obj"]


def main(ref: Path) -> None:
    out = Path(__file__).resolve().parent / "reference_scene"
    (out / "scenes").mkdir(parents=True, exist_ok=True)
    (out / "models").mkdir(parents=True, exist_ok=True)
    shutil.copyfile(ref / "scenes" / "obj-model.txt", out / "scenes" / "obj-model.txt")
    for m in MODELS:
        shutil.copyfile(ref / "models" / m, out / "models" / m)


if __name__ == "__main__":
    main(Path(sys.argv[1]))
