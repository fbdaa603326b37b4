"""Integer rewrites in the sampling and cell kernels (sesmesh_amd/csrc/
dev_sample.hpp, dev_cells.hpp), checked exhaustively against the
definitions they replace for every brick size the device path accepts
(even, 4..32):

* sample_owner's sub-block range per axis, min(v / sub, 1) and the lower
  bound on a shared face (crates/sesmesh/src/field/kernel.rs:495-532),
  rewritten without division for v in [0, 2 sub];
* k_assign_uniform's split of a sample index of the (b+1)^3 box into
  (x, y, z) by multiply-high with ceil(2^32 / d) instead of division;
* k_edges' and k_cell_emit's split of a cell index (< b^3) and of an edge
  id (< 3 b^3, cell = id / 3) likewise, and the sub-block of an edge's lower
  corner, min(v / sub, 1), as v >= sub.
"""


def test_owner_range_without_division():
    for b in range(4, 33, 2):
        sub = b // 2
        for v in range(b + 1):
            hi = min(v // sub, 1)
            lo = min(v // sub - 1, 1) if v > 0 and v % sub == 0 else hi
            assert (1 if v >= sub else 0) == hi
            assert (0 if v == sub else hi) == lo


def test_box_index_split_by_multiply_high():
    for b in range(4, 33, 2):
        side = b + 1
        plane = side * side
        m1 = ((1 << 32) + side - 1) // side
        m2 = ((1 << 32) + plane - 1) // plane
        for i in range(side ** 3):
            z = (i * m2) >> 32
            r = i - z * plane
            y = (r * m1) >> 32
            assert (r - y * side, y, z) == (i % side, (i // side) % side, i // plane)


def test_cell_index_split_by_multiply_high():
    for b in range(4, 33, 2):
        sub = b // 2
        m1 = ((1 << 32) + b - 1) // b
        m2 = ((1 << 32) + b * b - 1) // (b * b)
        for c in range(b ** 3):
            z = (c * m2) >> 32
            r = c - z * b * b
            y = (r * m1) >> 32
            x = r - y * b
            assert (x, y, z) == (c % b, (c // b) % b, c // (b * b))
            for v in (x, y, z):
                assert (1 if v >= sub else 0) == min(v // sub, 1)
