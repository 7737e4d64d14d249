"""Copies the glTF scenes the reference ships (data/GLTFs/ of the reference
repository: the PBR Cornell box and the nested-dielectrics scene, with their
binary buffers) into this folder, unchanged, as fixtures of the glTF ingest
tests. Run where the reference sources are present:
    python tests/golden/gltf/copy_fixtures.py /path/to/reference
"""
import os
import shutil
import sys

FILES = ["cornell_pbr.gltf", "cornell_pbr.bin", "nested-dielectrics.gltf", "nested-dielectrics.bin"]


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        shutil.copyfile(os.path.join(ref_root, "data", "GLTFs", f), os.path.join(here, f))
        print(f, os.path.getsize(os.path.join(here, f)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
