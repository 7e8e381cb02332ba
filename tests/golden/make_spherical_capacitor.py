"""Copies the reference's spherical-capacitor triangle meshes (scales 1-3) into gzip fixtures.

Source: KEMField/data/input_files/sphericalCapacitorFiles/SphericalCapacitor_{1,2,3}_triangles.dat
(read by ComputeSphericalCapacitor.cc:376-384 through ReadInTriangles, :1403-1476).  The files
are data (one triangle per row: P0, N1, N2, A, B, potential, permittivities above / below), kept
byte for byte.  Run in a tree that holds the reference at /root/reference:

    python tests/golden/make_spherical_capacitor.py
"""
import gzip
import os
import shutil

SRC = "/root/reference/KEMField/data/input_files/sphericalCapacitorFiles"
HERE = os.path.dirname(os.path.abspath(__file__))

for scale in (1, 2, 3):
    src = os.path.join(SRC, f"SphericalCapacitor_{scale}_triangles.dat")
    dst = os.path.join(HERE, f"spherical_capacitor_{scale}.dat.gz")
    with open(src, "rb") as f, gzip.GzipFile(dst, "wb", mtime=0) as g:
        shutil.copyfileobj(f, g)
    print(dst, os.path.getsize(dst))
