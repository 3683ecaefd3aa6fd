"""Copies the reference's data fixtures for the explicit step into tests/golden/.

Reads (never imports or runs) these data files of the reference checkout:
  examples/bowl.msh        Gmsh 4.1 mesh, 7555 nodes, 14788 triangles
  notebooks/basic.msh      Gmsh 4.2 mesh, 160 nodes, four named walls
  notebooks/topology.dat   tables written from bowl.msh by examples/Main.cpp:146-165
  notebooks/geometry.dat   node x y z written by examples/Main.cpp:136-144
  notebooks/out0.dat       w hu hv before one Euler step (examples/Main.cpp:192)
  notebooks/out1.dat       w hu hv after one Euler step, dt = 1e-3,
                           HLL<Einfeldt>, flat bottom (examples/Main.cpp:172-195)
and the 1-triangle Topology example of notebooks/Untitled1.ipynb.
The data are MIT-licensed (reference LICENSE.txt); see LICENSE-reference.txt.
"""
import gzip
import json
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

FILES = ["examples/bowl.msh", "notebooks/basic.msh", "notebooks/topology.dat", "notebooks/geometry.dat",
         "notebooks/out0.dat", "notebooks/out1.dat"]

for rel in FILES:
    src = os.path.join(REF, rel)
    dst = os.path.join(OUT, os.path.basename(rel) + ".gz")
    with open(src, "rb") as f, open(dst, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as g:
            shutil.copyfileobj(f, g)
    print(dst, os.path.getsize(dst))

# notebooks/Untitled1.ipynb cells 2 and 6-9: Topology.create on one triangle;
# the reference prints 3, 44070616, 140721212595328 (dangling references);
# the expected sizes are 3 nodes, 3 edges, 1 element.
one = {"numNodes": 3, "edgeNodes": [[0, 1], [0, 2], [1, 2]], "edgeElements": [[0, -1], [0, -1], [0, -1]],
       "elementNodes": [[0, 1, 2]], "elementEdges": [[0, 1, 2]], "elementNeighbours": [[-1, -1, -1]],
       "expected": {"num_nodes": 3, "num_edges": 3, "num_elements": 1}}
with open(os.path.join(OUT, "one_triangle.json"), "w") as f:
    json.dump(one, f, indent=1)
