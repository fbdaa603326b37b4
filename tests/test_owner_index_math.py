"""Integer rewrites in the sampling kernels (sesmesh_amd/csrc/dev_sample.hpp),
checked exhaustively against the definitions they replace for every brick
size the device path accepts (even, 4..32):

* sample_owner's sub-block range per axis, min(v / sub, 1) and the lower
  bound on a shared face (crates/sesmesh/src/field/kernel.rs:495-532),
  rewritten without division for v in [0, 2 sub];
* k_assign_uniform's split of a sample index of the (b+1)^3 box into
  (x, y, z) by multiply-high with ceil(2^32 / d) instead of division.
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
