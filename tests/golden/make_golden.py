"""Re-creates the reference's unit-test inputs as committed fixtures.

Every input of the reference's tests is closed form or drawn from glibc
rand() with a stated seed, so it is reproduced here without importing or
running anything from the reference:
  lj64   -- tests/test_lj.cu:18-27: srand(123), L*(rand()/(float)RAND_MAX), L=8
  morton -- tests/test_morton.cu:12-20: fmodf(i*3.7f, L), fmodf(i*5.3f, L), fmodf(i*7.1f, L), L=10
  nve64  -- tests/test_integrator.cu:22-42: 4x4x4 lattice spacing L/4 + 0.1, L=8,
            velocities 0.1*(rand()/(float)RAND_MAX - 0.5) after srand(42)
  nhke   -- tests/test_nose_hoover.cu:373-375: srand(12345), 10000 draws of
            KE = 0.5*N*T*(0.5 + (rand()/(float)RAND_MAX)*1.5), N=100, T=1.5
mini.data is the reference's own fixture file (tests/fixtures/mini.data).
Run: python tests/golden/make_golden.py  (writes tests/golden/ref_inputs.npz)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
libc = ctypes.CDLL("libc.so.6")
libc.rand.restype = ctypes.c_int
F = np.float32
RAND_MAX = F(2147483647)


def frand():
    return F(libc.rand()) / RAND_MAX


def main():
    out = {}
    libc.srand(123)
    L = F(8.0)
    lj = np.zeros((64, 3), np.float32)
    for i in range(64):
        lj[i, 0] = L * frand()
        lj[i, 1] = L * frand()
        lj[i, 2] = L * frand()
    out["lj64_xyz"] = lj
    L = F(10.0)
    mo = np.zeros((64, 3), np.float32)
    for i in range(64):
        mo[i] = [np.fmod(F(i) * F(3.7), L), np.fmod(F(i) * F(5.3), L), np.fmod(F(i) * F(7.1), L)]
    out["morton64_xyz"] = mo
    L = F(8.0)
    sp = L / F(4.0)
    lat = np.zeros((64, 3), np.float32)
    k = 0
    for ix in range(4):
        for iy in range(4):
            for iz in range(4):
                lat[k] = [F(ix) * sp + F(0.1), F(iy) * sp + F(0.1), F(iz) * sp + F(0.1)]
                k += 1
    libc.srand(42)
    vel = np.zeros((64, 3), np.float32)
    for i in range(64):
        for d in range(3):
            vel[i, d] = F(0.1) * (frand() - F(0.5))
    out["nve64_xyz"] = lat
    out["nve64_vel"] = vel
    libc.srand(12345)
    N, T = 100, F(1.5)
    ke = np.zeros(10000, np.float32)
    for t in range(10000):
        ke[t] = F(0.5) * F(N) * T * (F(0.5) + frand() * F(1.5))
    out["nhke"] = ke
    np.savez(os.path.join(HERE, "ref_inputs.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
