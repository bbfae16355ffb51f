"""Writes the MovePole fixtures from the reference's own data files (run in
the build container, where /root/reference exists):

  SpHarmRotMats_p6_double.txt   copied as is (3010 values)
  all_rot_mats_6_lat013.npy      latitudes 0, 1 and 3 of
                                 precomputed/all_rot_mats_6_double.txt,
                                 [3][84][84] row-major (row = output point)
  dumbbell_6.npy                 the first shape of precomputed/shape_gallery_6.txt
                                 (the dumbbell), [3][84]
"""
import os
import shutil

import numpy as np

SRC = "/root/reference/precomputed"
HERE = os.path.dirname(os.path.abspath(__file__))

p, n = 6, 84
shutil.copy(os.path.join(SRC, "SpHarmRotMats_p6_double.txt"), HERE)
R = np.loadtxt(os.path.join(SRC, "all_rot_mats_6_double.txt")).reshape(p + 1, n, n)
R = np.transpose(R, (0, 2, 1))                 # file blocks are column-major
np.save(os.path.join(HERE, "all_rot_mats_6_lat013.npy"), np.ascontiguousarray(R[[0, 1, 3]]))
vals, name = [], None
for line in open(os.path.join(SRC, "shape_gallery_6.txt")):
    line = line.strip()
    if line.startswith("#"):
        if name is not None:
            break
        name = line[1:].strip()
        continue
    if line:
        vals.append(float(line))
assert name == "dumbbell" and len(vals) == 3 * n
np.save(os.path.join(HERE, "dumbbell_6.npy"), np.array(vals).reshape(3, n))
