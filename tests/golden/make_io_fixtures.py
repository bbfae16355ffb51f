"""Writes the input-format fixtures from the reference's own data files (run
in the build container, where /root/reference exists):

  shape_gallery_6_head.txt   the first two shapes (dumbbell, oblate) of
                             precomputed/shape_gallery_6.txt, comments kept
  geometry_spec_head.txt     the first 6 lines of
                             precomputed/geometry_spec_sed_rand_512.txt
  shape_gallery_16_shape0.txt  the first shape of precomputed/shape_gallery_16.txt
                             (the only one the polydisperse sedimentation
                             suspension uses), comments kept
  geometry_spec_polysed_512.txt, vesprops_rand_polysed_512.txt
                             the reference's polydisperse sedimentation
                             suspension (experiment/polydisperse_config.in),
                             copied whole
"""
import shutil
import os

SRC = "/root/reference/precomputed"
HERE = os.path.dirname(os.path.abspath(__file__))

out, shapes = [], 0
for line in open(os.path.join(SRC, "shape_gallery_6.txt")):
    if line.startswith("#"):
        shapes += 1
        if shapes > 2:
            break
    out.append(line)
open(os.path.join(HERE, "shape_gallery_6_head.txt"), "w").writelines(out)
lines = open(os.path.join(SRC, "geometry_spec_sed_rand_512.txt")).readlines()[:6]
open(os.path.join(HERE, "geometry_spec_head.txt"), "w").writelines(lines)

out, shapes = [], 0
for line in open(os.path.join(SRC, "shape_gallery_16.txt")):
    if line.startswith("#"):
        shapes += 1
        if shapes > 1:
            break
    out.append(line)
open(os.path.join(HERE, "shape_gallery_16_shape0.txt"), "w").writelines(out)
for fn in ("geometry_spec_polysed_512.txt", "vesprops_rand_polysed_512.txt"):
    shutil.copyfile(os.path.join(SRC, fn), os.path.join(HERE, fn))
