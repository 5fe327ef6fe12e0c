"""The reference's own scene as a fixture: ref_tshirt_female.npz.

Reads the reference's resources (resources/cloth/tshirt2/tshirt2.obj and
resources/pose/female.obj under /root/reference) through the library's OBJ
reader and applies the set-up of the reference's main.cpp:24-32 (cloth
rotated 90 degrees about x and -4 about z, scaled 3.33, moved by (0, 1.98,
0.02); body scaled 3.226, moved by (0, 1.8, 0)).  Stores the cloth vertices
and faces and the body vertices, faces and per-vertex normals (Mesh::unified),
so the GPU tests and the bench can run that scene where the reference is not
present.  tests/test_obj.py checks that a fresh build equals the fixture when
the reference is present.

  python tests/golden/make_reference_meshes.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

REF = "/root/reference/resources"
CLOTH_OBJ = os.path.join(REF, "cloth", "tshirt2", "tshirt2.obj")
BODY_OBJ = os.path.join(REF, "pose", "female.obj")
OUT = os.path.join(HERE, "ref_tshirt_female.npz")


def build():
    from simulator_amd import scene as S
    sc = S.reference_scene(CLOTH_OBJ, BODY_OBJ)
    return dict(cloth_verts=sc.pos, cloth_faces=sc.faces, body_verts=sc.body.verts, body_faces=sc.body.tris,
                body_normals=sc.body.normals)


if __name__ == "__main__":
    d = build()
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, {k: v.shape for k, v in d.items()})
