"""Copy a few of the reference's bundled scene files into tests/golden/scenes/
as gzip-compressed fixtures (data only: .scn properties and PLY meshes).

    python tests/golden/make_scene_fixtures.py [/root/reference]

The tests unpack them into a temporary directory (tests/scene_fixtures.py).
"""
import gzip
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "bigmonkey": ["bigmonkey.scn", "bigmonkey-instances.scn", "bigmonkey-motion.scn", "bigmonkey.ply",
                  "room.ply", "bigmonkey-lights.ply"],
    "lightinstances": ["scene.scn", "mesh-00000.ply", "mesh-00001.ply", "mesh-00002.ply"],
    "simple": ["simple.scn", "simple-lights.ply", "simple-mat-cube1.ply", "simple-mat-cube2.ply",
               "simple-mat-back.ply", "simple-mat-back2.ply"],
}


def main(ref):
    for scene, names in FILES.items():
        out = os.path.join(HERE, "scenes", scene)
        os.makedirs(out, exist_ok=True)
        for n in names:
            with open(os.path.join(ref, "scenes", scene, n), "rb") as f:
                data = f.read()
            with open(os.path.join(out, n + ".gz"), "wb") as f:
                # mtime 0: the same input gives the same bytes
                with gzip.GzipFile(filename="", mode="wb", fileobj=f, compresslevel=9, mtime=0) as g:
                    g.write(data)
            print(scene, n, len(data))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
