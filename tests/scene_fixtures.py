"""Unpack the gzip scene fixtures of tests/golden/scenes/ (copied from the
reference's bundled scenes by tests/golden/make_scene_fixtures.py)."""
import gzip
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
SCENES = {
    "bigmonkey": ["bigmonkey.scn", "bigmonkey-instances.scn", "bigmonkey-motion.scn"],
    "lightinstances": ["scene.scn"],
    "simple": ["simple.scn"],
}


def unpack(scene, dest):
    """Write the scene's files, uncompressed, under dest/<scene>/; returns that directory."""
    out = os.path.join(str(dest), scene)
    os.makedirs(out, exist_ok=True)
    src = os.path.join(GOLDEN, scene)
    for n in sorted(os.listdir(src)):
        if not n.endswith(".gz"):
            continue
        target = os.path.join(out, n[:-3])
        if not os.path.exists(target):
            with gzip.open(os.path.join(src, n), "rb") as f, open(target, "wb") as g:
                g.write(f.read())
    return out
