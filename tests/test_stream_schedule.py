"""The register-streaming RK4 step (rk4_stream_kernel in rk4.hip) evaluates stage s
of row t-s at step t, keeps the loaded rows in an 8-slot ring indexed by the row mod 8
(each loaded 4 steps ahead) and the stage inputs in 4-slot rings indexed by the step
mod 4, pads the steps to a multiple of 8 (padding steps store nothing; loads past the
last step repeat its row), takes column neighbours from the adjacent lanes (whole-wave
DPP shifts; the edge lane reads 0) and covers the lattice with 56-column strips and
bands of min(R, L - 8) rows (R = 104, or 56 for small batches).  This emulates exactly that schedule with one numpy vector per wave (64 lanes),
in IEEE double with no fused operations, and checks one step against the oracle bit for
bit, plus that every site is written once and every load stays inside the lattice.
It pins the schedule; the GPU tests pin the kernel itself (test_gpu_rk4.py).
"""
import numpy as np
import pytest

HALO, COLS = 4, 56


def _f(p, c, up, dn, lf, rt):
    h = [np.zeros(64) for _ in range(3)]
    for n in (up, dn, lf, rt):                      # row+1, row-1, col-1, col+1
        for d in range(3):
            h[d] = h[d] + n[d] * p["J"]
    h[2] = h[2] + (2.0 * p["D"]) * c[2]
    h[2] = h[2] + p["Bz"]
    sx, sy, sz = c
    hx, hy, hz = h
    px, py, pz = hy * sz - hz * sy, hz * sx - hx * sz, hx * sy - hy * sx
    qx, qy, qz = sy * pz - sz * py, sz * px - sx * pz, sx * py - sy * px
    a, inv = p["alpha"] / p["len"], 1.0 / (1.0 + p["alpha"] * p["alpha"])
    return [(px + a * qx) * inv, (py + a * qy) * inv, (pz + a * qz) * inv]


def _from_left(v):
    o = np.zeros(64)
    o[1:] = v[:-1]
    return o


def _from_right(v):
    o = np.zeros(64)
    o[:-1] = v[1:]
    return o


def _k(p, c, up, dn):
    f = _f(p, c, up, dn, [_from_left(x) for x in c], [_from_right(x) for x in c])
    return [x * p["dt"] for x in f]


def stream_step(p, L, S_in, band_rows):
    out = np.full((L * L, 3), np.nan)
    writes = np.zeros(L * L, int)
    R = min(band_rows, L - 8)
    lane = np.arange(64)
    for band in range((L + R - 1) // R):
        for strip in range((L + COLS - 1) // COLS):
            c = strip * COLS - HALO + lane
            gc = ((c % L) + L) % L
            ocol = (lane >= HALO) & (lane < 64 - HALO) & (c < L)
            r0 = band * R
            rows = min(R, L - r0)

            def load(t):
                y = r0 - HALO + t
                y += L if y < 0 else 0
                y -= L if y >= L else 0
                assert 0 <= y < L
                return [S_in[y * L + gc, d].copy() for d in range(3)]

            z = lambda: [np.zeros(64) for _ in range(3)]
            Y2, Y3, Y4, A = ([z() for _ in range(4)] for _ in range(4))
            B = [z() for _ in range(8)]
            nstep = rows + 2 * HALO
            nstep8 = (nstep + 7) & ~7
            PF = 4
            for st in range(PF):
                B[st] = load(st)
            for t in range(nstep8):
                u = t % 8
                b0, b1, b2, b3 = u, (u + 7) & 7, (u + 6) & 7, (u + 5) & 7
                i0, i1, i2, i3 = u & 3, (u + 3) & 3, (u + 2) & 3, (u + 1) & 3
                i4, i5 = i0, i1
                B[(u + PF) & 7] = load(min(t + PF, nstep - 1))   # the slot of row t-4
                k = _k(p, B[b1], B[b0], B[b2])
                A[i1] = [B[b1][d] + k[d] / 6.0 for d in range(3)]
                Y2[i1] = [B[b1][d] + k[d] / 2.0 for d in range(3)]
                k = _k(p, Y2[i2], Y2[i1], Y2[i3])
                A[i2] = [A[i2][d] + k[d] / 3.0 for d in range(3)]
                Y3[i2] = [B[b2][d] + k[d] / 2.0 for d in range(3)]
                k = _k(p, Y3[i3], Y3[i2], Y3[i4])
                A[i3] = [A[i3][d] + k[d] / 3.0 for d in range(3)]
                Y4[i3] = [B[b3][d] + k[d] for d in range(3)]
                k = _k(p, Y4[i4], Y4[i3], Y4[i5])
                A[i4] = [A[i4][d] + k[d] / 6.0 for d in range(3)]
                orow = t - 2 * HALO
                if 0 <= orow < rows:
                    ax, ay, az = A[i4]
                    nrm = np.sqrt(ax * ax + ay * ay + az * az)
                    idx = (r0 + orow) * L + gc
                    for l in np.nonzero(ocol)[0]:
                        out[idx[l]] = [p["len"] * ax[l] / nrm[l], p["len"] * ay[l] / nrm[l],
                                       p["len"] * az[l] / nrm[l]]
                        writes[idx[l]] += 1
    return out, writes


@pytest.mark.parametrize("band_rows", [104, 56])
@pytest.mark.parametrize("L,kw", [(64, dict(D=0.1, Bz=0.1, alpha=0.01)),
                                  (100, dict(J=-0.7, D=0.05, Bz=0.2, alpha=0.2, spin_length=1.5)),
                                  (130, dict(D=0.1))])   # ragged last bands: 26 / 18 rows
def test_streaming_schedule_reproduces_oracle_step(orc, L, kw, band_rows):
    p = dict(J=1.0, D=0.0, Bz=0.0, alpha=0.0, dt=0.01, len=1.0)
    for key, v in kw.items():
        p["len" if key == "spin_length" else key] = v
    S0 = orc.init_spins(11, L * L) * p["len"]
    ref, _ = orc.run(orc.params(dt=0.01, **kw), L, S0, 1, 0, 0)
    got, writes = stream_step(p, L, S0, band_rows)
    assert (writes == 1).all()
    assert np.array_equal(got, ref)
