"""Copies the reference's committed data files used as golden fixtures into tests/golden/.

Run once in a container where the reference checkout is mounted at /root/reference.  The files
are data (scenes and the outputs of the reference's committed main() run), not source:
  data/stl_files/*.stl  (teapot.stl gzip-compressed)    -> tests/golden/stl/
  data/RX_*.dat                                         -> tests/golden/rx/
  data/output/{ElectricField.csv, RayPath.vtk, RxLocations.vtk, Icosahedron.vtk, Faces.vtk,
               Edges.vtk, BVH.vtk}                        -> tests/golden/output/
"""
import gzip
import os
import shutil

SRC = "/root/reference/SBRSolver_rev2/data"
DST = os.path.dirname(os.path.abspath(__file__))

def main():
    os.makedirs(os.path.join(DST, "stl"), exist_ok=True)
    os.makedirs(os.path.join(DST, "output"), exist_ok=True)
    os.makedirs(os.path.join(DST, "rx"), exist_ok=True)
    for f in sorted(os.listdir(os.path.join(SRC, "stl_files"))):
        if f == "teapot.stl":
            with open(os.path.join(SRC, "stl_files", f), "rb") as fi, \
                    gzip.GzipFile(os.path.join(DST, "stl", f + ".gz"), "wb", mtime=0) as fo:
                shutil.copyfileobj(fi, fo)
        elif f.endswith(".stl"):
            shutil.copyfile(os.path.join(SRC, "stl_files", f), os.path.join(DST, "stl", f))
    for f in sorted(os.listdir(SRC)):
        if f.startswith("RX_") and f.endswith(".dat"):
            shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, "rx", f))
    for f in ["ElectricField.csv", "RayPath.vtk", "RxLocations.vtk", "Icosahedron.vtk", "Faces.vtk", "Edges.vtk",
              "BVH.vtk"]:
        shutil.copyfile(os.path.join(SRC, "output", f), os.path.join(DST, "output", f))

if __name__ == "__main__":
    main()
