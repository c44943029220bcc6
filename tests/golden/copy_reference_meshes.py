"""Copy the mesh-file fixtures the reference's own tests use (test/cubeHex.msh:
a 6^3 hexahedral mesh of [-1,1]^3, Gmsh 2.2 ASCII, boundary quadrilaterals
tagged 1) into this folder. The copies are data, committed beside this script;
the GPU tests read the copies."""
import os
import shutil
import sys

SRC = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    for name in ("cubeHex.msh",):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(HERE, name))
        print("copied", name)
    sys.exit(0)
