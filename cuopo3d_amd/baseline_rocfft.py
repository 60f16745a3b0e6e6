"""The plain GPU comparator: library FFTs plus separate elementwise kernels.

This is the straightforward port the fused kernels are measured against: the
same operator sequence as the reference's SSFM (Solver3D.cuh:946-1011), with
every step a separate launch — torch.fft (rocFFT on ROCm) for the time and xy
transforms, torch elementwise ops for RK4, the table multiplies, the mirror
and the power.  It is a measurement baseline only; the product path is the
HIP kernels behind include/opo3d.h.
"""
import numpy as np
import torch


def _cis(phi):
    r = np.fmod(phi, 2 * np.pi)
    return np.cos(r) + 1j * np.sin(r)


def _signed(n):
    if n < 2:
        return np.zeros(1)
    j = np.arange(n)
    c = n // 2
    return ((j + c) % n) - c


class PlainSolve:
    """One solve on the GPU with rocFFT transforms and separate kernels."""

    def __init__(self, grid, p, fields, device="cuda"):
        self.nx, self.ny, self.nt = grid.nx, grid.ny, grid.nt
        self.p = p
        dev = torch.device(device)
        shape = (self.nt, self.ny, self.nx)
        self.api = torch.from_numpy(np.ascontiguousarray(fields[0]).reshape(shape)).to(dev)
        self.A = [torch.from_numpy(np.ascontiguousarray(f).reshape(shape)).to(dev).clone()
                  for f in fields]
        nt, nx, ny = self.nt, self.nx, self.ny
        w = np.zeros(nt)
        if nt > 1:
            F = -0.5 * nt / p.t_window + np.arange(nt) * (nt / p.t_window) / (nt - 1)
            c = nt // 2
            w[c:] = F[:c]
            w[:c] = F[c:]
            w *= 2 * np.pi
        dfx = 1.0 / (p.dx * nx) if nx > 1 else 0.0
        dfy = 1.0 / (p.dy * ny) if ny > 1 else 0.0
        fx, fy = _signed(nx), _signed(ny)
        q = (dfy * fy[:, None]) ** 2 + (dfx * fx[None, :]) ** 2
        to = lambda a: torch.from_numpy(a.astype(np.complex64)).to(dev)
        self.eiLz = [to((_cis(p.dz * w * (p.delta[f] + 0.5 * w * p.beta2[f]))
                         * np.exp(-0.5 * p.alpha[f] * p.dz)).reshape(nt, 1, 1)) for f in range(3)]
        self.eiQz = [to((_cis(-p.dz * 2 * np.pi ** 2 / p.kdiff[f] * q)
                         * np.exp(-0.5 * p.alpha[f] * p.dz))[None]) for f in range(3)]
        self.eiQfs = [to(_cis(p.fs_length * 2 * np.pi ** 2 / p.kdiff[1 + f] * q)[None]) for f in range(2)]
        self.egdd = to(_cis(0.5 * w * w * p.gdd).reshape(nt, 1, 1))
        self.fb = [complex(p.fb_re[f], p.fb_im[f]) for f in range(2)]
        self.kap = [float(k) for k in p.kappa]
        self.trip = 0

    def _rhs(self, ap, as_, ai):
        kp, ks, ki = self.kap
        return (1j * kp) * as_ * ai, (1j * ks) * ap * ai.conj(), (1j * ki) * ap * as_.conj()

    def rk4(self):
        h = 0.5 * self.p.dz
        c1, c3 = (0.5 * h, h) if self.p.rk4_h_in_stages else (0.5, 1.0)
        A = self.A
        k1 = self._rhs(*A)
        k2 = self._rhs(*[a + c1 * k for a, k in zip(A, k1)])
        k3 = self._rhs(*[a + c1 * k for a, k in zip(A, k2)])
        k4 = self._rhs(*[a + c3 * k for a, k in zip(A, k3)])
        self.A = [a + (x1 + 2 * x2 + 2 * x3 + x4) * h / 6
                  for a, x1, x2, x3, x4 in zip(A, k1, k2, k3, k4)]

    def dispersion(self):
        self.A = [torch.fft.fft(torch.fft.ifft(a, dim=0) * t, dim=0)
                  for a, t in zip(self.A, self.eiLz)]

    def diffraction(self):
        self.A = [torch.fft.ifft2(torch.fft.fft2(a, dim=(1, 2)) * t, dim=(1, 2))
                  for a, t in zip(self.A, self.eiQz)]

    def ssfm(self):
        for _ in range(self.p.n_slices):
            self.rk4()
            self.dispersion()
            self.diffraction()
            self.rk4()

    def round_trip(self):
        p = self.p
        self.A[0] = self.api.clone()
        self.ssfm()
        if p.fs_on and not (p.skip_fs_on_last_trip and self.trip == p.n_round_trips - 1):
            for f in range(2):
                self.A[1 + f] = torch.fft.ifft2(torch.fft.fft2(self.A[1 + f], dim=(1, 2)) * self.eiQfs[f],
                                                dim=(1, 2))
        if p.gdd_on:
            for f in range(2):
                self.A[1 + f] = torch.fft.fft(torch.fft.ifft(self.A[1 + f], dim=0) * self.egdd, dim=0)
        for f in range(2):
            self.A[1 + f] = self.A[1 + f] * self.fb[f]
        self.trip += 1
        # device tensor (3,): no host synchronisation inside the trip loop
        return torch.stack([float(p.power_prefactor[f]) * (a.real.double() ** 2 + a.imag.double() ** 2).sum()
                            for f, a in enumerate(self.A)])

    def fields(self):
        return [a.cpu().numpy() for a in self.A]


def main():
    """Time the comparator on the 64x64x256 cavity workload; print one JSON line.

    Run as its own process: torch brings its own HIP runtime, which cannot
    share a process with the system runtime libopo3d.so is linked against."""
    import argparse
    import json
    import time
    from . import configs
    ap = argparse.ArgumentParser()
    ap.add_argument("--trips", type=int, default=4)
    ap.add_argument("--sweep-solves", type=int, default=0,
                    help="time the first N solves of the mixed-size sweep one at a time instead")
    args = ap.parse_args()
    if args.sweep_solves > 0:
        return sweep_main(args.sweep_solves, args.trips)
    g, p, _, f = configs.cavity_64(n_round_trips=args.trips + 1)
    ps = PlainSolve(g, p, f)
    ps.round_trip()  # warm-up: FFT plans and caches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.trips):
        ps.round_trip()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cells = g.nx * g.ny * g.nt
    print(json.dumps({"value": cells * p.n_slices * args.trips / dt, "unit": "cell-steps/s",
                      "trips": args.trips, "ms_per_trip": dt / args.trips * 1e3,
                      "what": "torch.fft (rocFFT) transforms + torch elementwise RK4 and "
                              "multiplies, same operator sequence, separate launches"}))


def sweep_main(n_solves, trips):
    """The plain path one solve at a time over the first solves of the
    mixed-size sweep (the same draws as the batched sweep)."""
    import json
    import time
    from . import configs
    draws = configs.sweep_draws(n_solves=64, seed=3)[:n_solves]
    work = 0
    tot = 0.0
    for d in draws:
        g, p, _, f = configs.sweep_solve(d, n_round_trips=trips + 1)
        ps = PlainSolve(g, p, f)
        ps.round_trip()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(trips):
            ps.round_trip()
        torch.cuda.synchronize()
        tot += time.perf_counter() - t0
        work += g.nx * g.ny * g.nt * p.n_slices * trips
        del ps
    print(json.dumps({"value": work / tot, "unit": "cell-steps/s", "solves": n_solves,
                      "trips": trips, "seconds": tot}))


if __name__ == "__main__":
    main()
