"""Writes morton_d3_b1_e1.json: the initial-grid cells of kdgrid.cu:141-164 for
D = 3, bitsPerDim = 1, extraImgBits = 1, derived by hand from the bit layout:

  k = 0 < b      : xq_d bit 0 <- index bit D*0 + (D-1-d)  -> x: bit 2, y: bit 1, t: bit 0
  k = 1 in [b,b+e): xq_d bit 1 <- index bit D*1 + 2*0 + 1 - d -> x: bit 4, y: bit 3

so leaf i = (b4 b3 b2 b1 b0) has x = b4 b2, y = b3 b1, t = b0 (e.g. leaf 22 =
10110b -> x = 3, y = 1, t = 0).
"""
import json
import os

cells = []
for i in range(32):
    b = [(i >> k) & 1 for k in range(5)]
    cells.append({"leaf": i, "xq": [b[2] | (b[4] << 1), b[1] | (b[3] << 1), b[0]]})
assert cells[22]["xq"] == [3, 1, 0]
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "morton_d3_b1_e1.json"), "w") as f:
    json.dump({"dim": 3, "bits_per_dim": 1, "extra_img_bits": 1, "cells": cells}, f, indent=0)
