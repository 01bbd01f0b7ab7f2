"""The CPU restatement of the ray coherence engine (oracle/coherence.py), pinned
by values derived by hand from src/gpgpu-sim/ray_coherency_engine.cc and by the
reference's own greedy scheduling loop."""
import numpy as np

import coherence


def test_hash_comp_hand_derived():
    # 1.0f = 0x3F800000: sign 0, top exponent bits (>>29)&3 = 1, top mantissa bits (>>21)&3 = 0 -> 0b00100
    # -0.75f = 0xBF400000: sign 1, (>>29)&3 = 1, (>>21)&3 = 2 -> 0b10110
    assert coherence.hash_comp(np.float32([1.0, -0.75]), 2).tolist() == [4, 22]


def test_direction_hash_hand_derived(rtx):
    # sphere bits 2: theta = acos(z)/pi*180 >> 6, phi = (atan2(y, x) + pi)/pi*180 >> 6
    # (0,0,1): theta 0, phi 180 -> (2 << 2) | 0 = 8;  (1,0,0): theta 90 -> 1, phi 180 -> 9;
    # (-1,0,0): phi 360 -> 5 -> (5 << 2) | 1 = 21
    r = np.zeros(3, rtx.RAY_DTYPE)
    r["dir"] = [(0, 0, 1), (1, 0, 0), (-1, 0, 0)]
    assert coherence.ray_hash(r, "d", sphere_bits=2).tolist() == [8, 9, 21]


def test_origin_grid_hand_derived(rtx):
    # world [0,8]^3, 3 bits: cell = clamp((o - 0) / 8 * 8, 0, 7); (1.5, 7.9, 9) -> (1, 7, 7); (-2, 0, 4) -> (0, 0, 4)
    o = np.float32([[1.5, 7.9, 9.0], [-2.0, 0.0, 4.0]])
    h = coherence.hash_origin_grid(o, 3, (0, 0, 0), (8, 8, 8))
    assert h.tolist() == [(1 << 6) | (7 << 3) | 7, (0 << 6) | (0 << 3) | 4]


def test_packet_schedule_hand_derived():
    # packet 2, hashes [5,3,5,5,3,7]: pools 3:[1,4] 5:[0,2,3] 7:[5]; largest first: 5 -> [0,2],
    # then 3 (2 left) -> [1,4], then 5 and 7 tie at 1 -> lowest hash 5 -> [3], then [5]
    assert coherence.packet_order([5, 3, 5, 5, 3, 7], 2).tolist() == [0, 2, 1, 4, 3, 5]


def test_greedy_schedule_equals_the_chunk_sort():
    """The reference takes the largest pool each time (ray_coherency_engine.cc:155-193);
    sorting chunks by rays left then hash gives the same order."""
    rng = np.random.default_rng(3)
    for n, span, packet in ((5000, 40, 64), (3000, 7, 32), (777, 500, 1), (4096, 3, 5)):
        h = rng.integers(0, span, n).astype(np.uint64)
        assert np.array_equal(coherence.packet_order(h, packet), coherence.packet_order_sorted(h, packet))
