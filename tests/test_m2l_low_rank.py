"""Host logic of the low-rank M2L tables (gpurify_amd/csrc/gp_m2l_svd.cpp),
compiled here with g++ and checked against numpy's symmetric eigensolver on
the interaction matrices of the Chebyshev bbFMM (the uncompressed M2L
definition of fmm/bbfmm.rs:156-166)."""
import itertools
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpurify_amd", "csrc")

HARNESS = r"""
#include "gp_m2l_svd.h"
#include <cstdio>
#include <cstdlib>
#include <vector>
int main(int argc, char **argv) {
  const int mode = atoi(argv[1]);
  FILE *f = fopen(argv[2], "rb");
  int n3, ld, kr;
  if (fread(&n3, 4, 1, f) != 1 || fread(&ld, 4, 1, f) != 1 || fread(&kr, 4, 1, f) != 1) return 2;
  const size_t cnt = mode == 0 ? (size_t)343 * ld * ld : (size_t)n3 * n3;
  std::vector<double> in(cnt);
  if (fread(in.data(), 8, cnt, f) != cnt) return 2;
  fclose(f);
  FILE *o = fopen(argv[3], "wb");
  if (mode == 0) {
    std::vector<double> U, Gc;
    gp::m2l_low_rank(n3, ld, in.data(), kr, 4, U, Gc);
    fwrite(U.data(), 8, U.size(), o);
    fwrite(Gc.data(), 8, Gc.size(), o);
  } else {
    std::vector<double> w, V;
    gp::sym_eig(n3, in, w, V);
    fwrite(w.data(), 8, w.size(), o);
    fwrite(V.data(), 8, V.size(), o);
  }
  fclose(o);
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("lowrank")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread", "-I", CSRC, str(src),
                    os.path.join(CSRC, "gp_m2l_svd.cpp"), "-o", str(exe)], check=True)
    return d, exe


def run(harness, mode, header, data, nout):
    d, exe = harness
    fin, fout = d / "in.bin", d / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array(header, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(data, dtype=np.float64).tobytes())
    subprocess.run([str(exe), str(mode), str(fin), str(fout)], check=True)
    out = np.fromfile(fout, dtype=np.float64)
    assert out.size == nout
    return out


def m2l_tables(n):
    x = np.cos((2 * np.arange(n) + 1) * np.pi / (2 * n))
    nodes = np.array(list(itertools.product(x, x, x)))  # m = a n^2 + b n + c
    n3 = n ** 3
    ld = (n3 + 31) // 32 * 32
    G = np.zeros((343, ld, ld))
    for o0, o1, o2 in itertools.product(range(-3, 4), repeat=3):
        if max(abs(o0), abs(o1), abs(o2)) <= 1:
            continue
        oi = (o0 + 3) * 49 + (o1 + 3) * 7 + (o2 + 3)
        d = nodes[:, None, :] - nodes[None, :, :] - 2.0 * np.array([o0, o1, o2])
        G[oi, :n3, :n3] = 1.0 / np.linalg.norm(d, axis=2)
    return G, n3, ld


def test_sym_eig_matches_numpy(harness):
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 40):
        A = rng.normal(size=(n, n))
        A = A + A.T
        out = run(harness, 1, (n, n, 0), A, n + n * n)
        w, V = out[:n], out[n:].reshape(n, n)
        assert np.allclose(w, np.sort(np.linalg.eigvalsh(A))[::-1], atol=1e-12 * np.abs(A).max())
        assert np.allclose(V.T @ V, np.eye(n), atol=1e-12)
        assert np.allclose(A @ V, V * w, atol=1e-11 * np.abs(A).max())
    # repeated eigenvalues: a rotated diagonal
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    lam = np.array([3.0, 3.0, 3.0, 1.0, 1.0, -2.0])
    out = run(harness, 1, (6, 6, 0), Q @ np.diag(lam) @ Q.T, 42)
    assert np.allclose(out[:6], lam, atol=1e-13)


@pytest.mark.parametrize("n,kr,bound", [(5, 64, 1e-4), (6, 64, 1e-4), (4, 32, 2e-3)])
def test_low_rank_tables(harness, n, kr, bound):
    G, n3, ld = m2l_tables(n)
    out = run(harness, 0, (n3, ld, kr), G, n3 * kr + 343 * kr * kr)
    U = out[:n3 * kr].reshape(n3, kr)
    Gc = out[n3 * kr:].reshape(343, kr, kr)
    assert np.allclose(U.T @ U, np.eye(kr), atol=1e-12)
    Gs = G[:, :n3, :n3]
    # the same subspace as numpy's leading eigenvectors of sum_o G_o G_o^T
    S = (Gs @ Gs.transpose(0, 2, 1)).sum(axis=0)
    wa, V = np.linalg.eigh(S)
    wa, Vk = wa[::-1], V[:, ::-1][:, :kr]
    # rounding of S (~eps ||S||) moves the subspace by ~eps ||S|| / gap
    cond = 2.2e-16 * wa[0] / (wa[kr - 1] - wa[kr])
    assert np.linalg.norm(U @ U.T - Vk @ Vk.T) < max(1e-9, 100 * cond)
    # Gc_o = U^T G_o U, zero on the near offsets
    ref = U.T @ Gs @ U
    assert np.abs(Gc - ref).max() < 1e-12 * np.abs(ref).max()
    near = [(o0 + 3) * 49 + (o1 + 3) * 7 + (o2 + 3)
            for o0, o1, o2 in itertools.product(range(-1, 2), repeat=3)]
    assert np.all(Gc[near] == 0.0)
    # projection error of every far matrix: small, and not below the
    # truncated spectrum (a wrong subspace or a transposed table fails)
    far = [o for o in range(343) if o not in near]
    err = [np.linalg.norm(Gs[o] - U @ Gc[o] @ U.T) / np.linalg.norm(Gs[o]) for o in far]
    assert max(err) < bound
    w = np.sort(np.linalg.eigvalsh(S))[::-1]
    total = sum(np.linalg.norm(Gs[o]) ** 2 for o in far)
    resid = sum(np.linalg.norm(Gs[o] - U @ Gc[o] @ U.T) ** 2 for o in far)
    # sum_o ||G_o - P G_o P||^2 >= sum_o ||(I - P) G_o||^2 = sum of the dropped eigenvalues
    assert resid >= 0.999 * w[kr:].sum() and resid < 2.01 * w[kr:].sum()
    assert resid / total < bound ** 2
