"""Digests of the oracle's hit records for the full plain workload: the
101,932-triangle scene (box + level-6 icosphere + 20,000 random triangles,
scenes.plain_scene(seed=1)) and its 1,048,576 rays (scenes.plain_rays(1 << 20,
seed=2)), traced by oracle.OracleBVH (BVHAccel::Intersect over the SAH tree).

    python tests/golden/make_plain_full_digest.py

Writes tests/golden/plain_full_digest.json: one SHA-1 per chunk of 4,096 rays
of the normalized hit records (a miss keeps t and its null indices; its
barycentrics count as zero).  The GPU test compares the device's chunks with
these and runs the oracle only on a chunk that differs, where the parity rule
(edge band, exact ties) decides.  The oracle takes about a minute here.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CHUNK = 4096
N_RAYS = 1 << 20


def main():
    import oracle
    from luxcore_amd import scenes
    from tests.parity import chunk_digests
    meshes = scenes.plain_scene(seed=1)
    rays = scenes.plain_rays(N_RAYS, seed=2)
    t0 = time.time()
    hits = oracle.OracleBVH(meshes).intersect(rays)
    out = {"scene": "scenes.plain_scene(seed=1)", "rays": "scenes.plain_rays(1 << 20, seed=2)",
           "triangles": int(sum(m.ntris for m in meshes)), "n_rays": N_RAYS, "chunk": CHUNK,
           "hits": int((hits["meshIndex"] != 0xFFFFFFFF).sum()), "sha1": chunk_digests(hits, CHUNK)}
    with open(os.path.join(HERE, "plain_full_digest.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"oracle {time.time() - t0:.1f}s, {out['hits']} hits, {len(out['sha1'])} chunks")


if __name__ == "__main__":
    main()
