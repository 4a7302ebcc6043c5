"""glibc's srand() and rand() per world (a23): the oracle against this machine's own C library draw for draw (the
TYPE_3 integer algorithm is the same in every glibc), and the library's host build against the oracle (states and
draws); the device build against the oracle is marked gpu."""

import ctypes
import ctypes.util

import numpy as np
import pytest
import torch

import rand_oracle as R
from exact_ballpark_amd import _native as N

SEEDS = [0, 1, 2, 3, 42, 127773, 2147483646, 2147483647, 2147483648, 2147483649, 4294967295, 123456789, 987654321]


def _libc_draws(seed: int, k: int) -> list:
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(ctypes.c_uint(seed))
    return [libc.rand() for _ in range(k)]


@pytest.mark.parametrize("seed", SEEDS + list(np.random.default_rng(7).integers(0, 2 ** 32, 5)))
def test_oracle_matches_libc(seed):
    seed = int(seed)
    g = R.Rand(seed)
    assert [g.next() for _ in range(2000)] == _libc_draws(seed, 2000)


def _seeds():
    rng = np.random.default_rng(8)
    return np.array(SEEDS + list(rng.integers(0, 2 ** 32, 500)), dtype=np.uint32)


def test_host_build_matches_oracle():
    L = N.cpu()
    seeds = _seeds()
    W, K = seeds.size, 64
    st = np.zeros((W, 33), dtype=np.int32)
    out = np.zeros((W, K), dtype=np.int32)
    assert L.xb_cpu_rand_seed(seeds.ctypes.data, W, st.ctypes.data) == 0
    for w, s in enumerate(seeds.tolist()):
        assert st[w].tolist() == R.Rand(s).state()
    assert L.xb_cpu_rand_draw(st.ctypes.data, W, K, out.ctypes.data) == 0
    for w, s in enumerate(seeds.tolist()):
        g = R.Rand(s)
        assert out[w].tolist() == [g.next() for _ in range(K)] and st[w].tolist() == g.state()


@pytest.mark.gpu
def test_device_matches_oracle():
    seeds = _seeds()
    W, K = seeds.size, 64
    st = torch.zeros(W, 33, dtype=torch.int32, device="cuda")
    out = torch.zeros(W, K, dtype=torch.int32, device="cuda")
    sd = torch.from_numpy(seeds.view(np.int32)).to("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    L = N.hip()
    N.check(L.xb_rand_seed(sd.data_ptr(), W, st.data_ptr(), stream), "xb_rand_seed")
    N.check(L.xb_rand_draw(st.data_ptr(), W, K, out.data_ptr(), stream), "xb_rand_draw")
    st, out = st.cpu().numpy(), out.cpu().numpy()
    for w, s in enumerate(seeds.tolist()):
        g = R.Rand(s)
        assert out[w].tolist() == [g.next() for _ in range(K)] and st[w].tolist() == g.state()
