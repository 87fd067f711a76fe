"""Learned absorbing boundary for Acoustic2D (FP32 interior, BF16-MFMA network).

pyfds has no absorbing boundary: its grids end in zero ghosts and the x stencil wraps from row
to row (``pyfds/fields.py:273-365``), so an open domain has to be emulated by padding the grid
until reflections arrive too late to matter.  Here a G x G grid (an N x N interior plus a
16-cell layer on every side) gets its layer from a small convolutional network after every
step instead:

* per side, two linear 1D convolutions along the boundary (kernel 5): 48 input channels (the
  8 interior rows next to the layer, p / vx / vy, current and previous step) -> 32 -> 48 output
  channels (the 16 layer rows of p / vx / vy); 4 sides x 30 720 weights = 61 440 parameters, no
  biases, so the map is linear and scale-equivariant like the wave equation it stands in for;
* velocities enter in pressure units (times rho0 c0) so all channels have one scale;
* weights are bf16, products accumulate in fp32 on the matrix cores
  (``pyfds_amd/csrc/fds_lbc.hip``); inference runs as one fused kernel per step inside the
  device step loop (``fds_set_learned_boundary``), captured in the same HIP graphs.

Training data come from the plain method on the enlarged grid (default 3072^2, the interior
centred): the layer and band rows are recorded every step while waves from random sources and
inclusions cross them, and the network is fit to predict the recorded layer from the recorded
band (teacher forcing) with the project's own forward, backward and Adam kernels.  The same
enlarged run judges the result: reflected energy = sum over the interior of
(p_learned - p_enlarged)^2 / sum p_enlarged^2 over the window after the wavefront reaches the
layer.  ``LearnedBoundary.run`` falls back to the plain enlarged-domain run when the validated
reflected energy is above the tolerance.
"""
import ctypes as ct
import time

import numpy as np

from . import _native as nat
from . import autotune
from .acoustics import Acoustic2D, AcousticMaterial
from . import regions

LAYER, BAND, CIN, HIDDEN, COUT, TAPS = 16, 8, 48, 32, 48, 5
DEPTH_REC = LAYER + BAND


def lbc_cell(side, d, a, g):
    """(side, depth, along) -> flat index of the G x G grid (same map as the kernels)."""
    d, a = np.asarray(d), np.asarray(a)
    if side == 0:
        return d * g + a
    if side == 1:
        return (g - 1 - d) * g + a
    if side == 2:
        return a * g + d
    return a * g + (g - 1 - d)


def ricker(t, f0):
    t0 = 1.5 / f0
    arg = (np.pi * f0 * (t - t0)) ** 2
    return (1 - 2 * arg) * np.exp(-arg)


class Scene:
    """One source and a few inclusions inside the interior (interior coordinates)."""

    def __init__(self, source, inclusions, f0):
        self.source, self.inclusions, self.f0 = source, inclusions, f0

    @classmethod
    def random(cls, rng, n_int, margin=48, f0=5e5):
        inc = []
        for _ in range(int(rng.integers(1, 5))):
            w, h = (int(v) for v in rng.integers(n_int // 16, n_int // 5, size=2))
            x0 = int(rng.integers(margin, n_int - margin - w))
            y0 = int(rng.integers(margin, n_int - margin - h))
            inc.append((x0, y0, w, h, float(rng.uniform(1000, 3000)),
                        float(rng.uniform(800, 3000))))
        # sources anywhere between the margin and the centre: every incidence angle occurs
        src = tuple(int(v) for v in rng.integers(margin, n_int - margin, size=2))
        return cls(src, inc, f0)


class LearnedBoundary:
    """Train, attach and judge the learned boundary layer (see module docstring)."""

    def __init__(self, n_int=1024, big=3072, dx=1e-4, c0=1500.0, rho0=1000.0, cmax=3000.0,
                 seed=0, device=0):
        import torch
        self.torch = torch
        self.lib = nat.load()
        self.n_int, self.big, self.dx = n_int, big, dx
        self.G = n_int + 2 * LAYER
        self.off = (big - self.G) // 2
        self.c0, self.rho0 = c0, rho0
        self.dt = 0.5 * dx / cmax
        self.cs = np.array([1.0, rho0 * c0, rho0 * c0], np.float32)
        self.dev = torch.device('cuda', device)
        g = torch.Generator().manual_seed(seed)
        self.w1 = (torch.randn(4, HIDDEN, TAPS, CIN, generator=g)
                   * (0.3 / np.sqrt(TAPS * CIN))).to(self.dev)
        self.w2 = (torch.randn(4, COUT, TAPS, HIDDEN, generator=g)
                   * (0.3 / np.sqrt(TAPS * HIDDEN))).to(self.dev)
        self.opt = [torch.zeros_like(w) for w in (self.w1, self.w1, self.w2, self.w2)]
        self.w1b = torch.empty(self.w1.shape, dtype=torch.int16, device=self.dev)
        self.w1t = torch.empty((4, CIN, TAPS, HIDDEN), dtype=torch.int16, device=self.dev)
        self.w2b = torch.empty(self.w2.shape, dtype=torch.int16, device=self.dev)
        self.w2t = torch.empty((4, HIDDEN, TAPS, COUT), dtype=torch.int16, device=self.dev)
        self.adam_step = 0
        self.frames = None
        self.validated = None
        self._pack()

    # ------------------------------------------------------------------ helpers
    @property
    def n_params(self):
        return self.w1.numel() + self.w2.numel()

    def _stream(self):
        return ct.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    @staticmethod
    def _p(t):
        return ct.c_void_p(t.data_ptr())

    def _pack(self):
        nat.check(self.lib.fds_lbc_pack(self._p(self.w1), self._p(self.w1b), self._p(self.w1t),
                                        4, HIDDEN, TAPS, CIN, self._stream()))
        nat.check(self.lib.fds_lbc_pack(self._p(self.w2), self._p(self.w2b), self._p(self.w2t),
                                        4, COUT, TAPS, HIDDEN, self._stream()))

    def field(self, scene, total, offset, steps, form='fp32'):
        """Acoustic2D of total^2 cells with the scene's interior cell (0, 0) at
        (offset, offset); an additive Ricker pressure source."""
        fld = Acoustic2D(x_samples=total, x_delta=self.dx, y_samples=total, y_delta=self.dx,
                         t_samples=steps, t_delta=self.dt,
                         material=AcousticMaterial(self.c0, self.rho0))
        fld.form = form
        base = LAYER + offset
        for x0, y0, w, h, c, rho in scene.inclusions:
            xs = np.arange(base + x0, base + x0 + w)
            ys = np.arange(base + y0, base + y0 + h)
            fld.add_material_region(regions.Region((xs[None, :] + ys[:, None] * total).ravel()),
                                    AcousticMaterial(c, rho))
        sx, sy = scene.source
        t = np.arange(steps + 1) * self.dt
        fld.pressure.add_boundary(regions.Region([(base + sy) * total + base + sx]),
                                  value=ricker(t, scene.f0), additive=True)
        return fld

    def frame_cells(self, total, offset):
        """Flat indices (in a total^2 grid) of the recorded frame: side, depth < 24, along."""
        a = np.arange(self.G)
        cells = []
        for s in range(4):
            for d in range(DEPTH_REC):
                loc = lbc_cell(s, d, a, self.G)
                iy, ix = np.divmod(loc, self.G)
                cells.append((iy + offset) * total + ix + offset)
        return np.concatenate(cells).astype(np.int64)

    def wave_arrival(self, scene, margin=0):
        """Step at which the direct wave reaches the layer (plus the pulse onset)."""
        sx, sy = scene.source
        dist = min(sx, sy, self.n_int - 1 - sx, self.n_int - 1 - sy) + margin
        return int(dist * self.dx / (self.c0 * self.dt))

    # ------------------------------------------------------------------ data
    def record(self, scene, steps, start, chunk=400):
        """Plain enlarged-domain run; returns frames [steps - start][3][4][24][G] float32 of
        steps start..steps-1 (end-of-step states)."""
        fld = self.field(scene, self.big, self.off, steps)
        fld.assemble_matrices()
        if start:
            fld.simulate(start)
        cells = self.frame_cells(self.big, self.off)
        runner = fld._runner()
        out = []
        while fld.step < steps:
            n = min(chunk, steps - fld.step)
            rec = runner.run_recording(fld, fld.step, n, [cells, cells, cells])
            out.append(np.stack([r.reshape(n, 4, DEPTH_REC, self.G) for r in rec], axis=1))
            fld.step += n
        runner.close()
        return np.concatenate(out)

    def generate(self, n_runs=3, seed=100, after=1200):
        """Record ``n_runs`` random scenes from shortly before the first arrival to ``after``
        steps past it; keeps the frames on the GPU."""
        rng = np.random.default_rng(seed)
        frames, self.run_len = [], []
        for _ in range(n_runs):
            sc = Scene.random(rng, self.n_int)
            t_hit = self.wave_arrival(sc)
            start = max(0, t_hit - 100)
            fr = self.record(sc, t_hit + after, start)
            frames.append(fr)
            self.run_len.append(fr.shape[0])
        T = min(self.run_len)
        arr = np.stack([f[:T] for f in frames])                 # [runs][T][3][4][D][G]
        self.T = T
        self.frames = self.torch.from_numpy(np.ascontiguousarray(arr)).to(self.dev)
        # p scale of band + inner layer rows per (run, step, side), for sample selection
        p = arr[:, :, 0, :, LAYER - 4:LAYER + BAND, LAYER:self.G - LAYER]
        self.rms = np.sqrt(np.mean(p.astype(np.float64) ** 2, axis=(3, 4)))   # [runs][T][4]
        return arr.shape

    # ------------------------------------------------------------------ training
    def _forward(self, x, side, B):
        t = self.torch
        P = self.G
        h = t.empty((B, P, HIDDEN), dtype=t.float32, device=self.dev)
        hb = t.empty((B, P, HIDDEN), dtype=t.int16, device=self.dev)
        y = t.empty((B, P, COUT), dtype=t.float32, device=self.dev)
        nat.check(self.lib.fds_lbc_conv_fwd(self._p(x), self._p(self.w1b), self._p(side),
                                            self._p(h), self._p(hb), B, P, CIN, HIDDEN, TAPS,
                                            self._stream()))
        nat.check(self.lib.fds_lbc_conv_fwd(self._p(hb), self._p(self.w2b), self._p(side),
                                            self._p(y), None, B, P, HIDDEN, COUT, TAPS,
                                            self._stream()))
        return hb, y

    def train(self, iters=1500, batch=32, lr=2e-3, seed=0, log_every=0):
        t = self.torch
        rng = np.random.default_rng(seed)
        runs, T, _ = self.rms.shape
        cand = [(r, s, side) for r in range(runs) for s in range(1, T) for side in range(4)
                if self.rms[r, s, side] > 0.02 * self.rms[r].max()]
        cand = np.array(cand, dtype=np.int64)
        B, P = batch, self.G
        x = t.empty((B, P, CIN), dtype=t.int16, device=self.dev)
        tgt = t.empty((B, P, COUT), dtype=t.float32, device=self.dev)
        dyb = t.empty((B, P, COUT), dtype=t.int16, device=self.dev)
        dh = t.empty((B, P, HIDDEN), dtype=t.float32, device=self.dev)
        dhb = t.empty((B, P, HIDDEN), dtype=t.int16, device=self.dev)
        part1 = t.empty((B, HIDDEN, TAPS, CIN), dtype=t.float32, device=self.dev)
        part2 = t.empty((B, COUT, TAPS, HIDDEN), dtype=t.float32, device=self.dev)
        g1 = t.empty_like(self.w1)
        g2 = t.empty_like(self.w2)
        wch = np.where(np.arange(COUT) % LAYER >= LAYER - 4, 1.0, 0.05).astype(np.float32)
        wch_d = t.from_numpy(wch).to(self.dev)
        nblk = 256
        loss_part = t.empty(nblk, dtype=t.float32, device=self.dev)
        cs = self.cs.ctypes.data_as(ct.c_void_p)
        history = []
        for it in range(iters):
            pick = cand[rng.integers(0, len(cand), size=B)]
            run = t.from_numpy(pick[:, 0].astype(np.int32)).to(self.dev)
            step = t.from_numpy(pick[:, 1].astype(np.int32)).to(self.dev)
            side = t.from_numpy(pick[:, 2].astype(np.int32)).to(self.dev)
            scale = t.from_numpy((1.0 / self.rms[pick[:, 0], pick[:, 1], pick[:, 2]])
                                 .astype(np.float32)).to(self.dev)
            lo = t.from_numpy(np.where(pick[:, 2] < 2, 0, LAYER).astype(np.int32)).to(self.dev)
            hi = t.from_numpy(np.where(pick[:, 2] < 2, P, P - LAYER).astype(np.int32)).to(self.dev)
            nat.check(self.lib.fds_lbc_gather(self._p(self.frames), self._p(run), self._p(step),
                                              self._p(side), self._p(scale), self._p(x),
                                              self._p(tgt), B, self.T, DEPTH_REC, P, LAYER, cs,
                                              self._stream()))
            hb, y = self._forward(x, side, B)
            norm = 1.0 / (B * P * 4.0)
            nat.check(self.lib.fds_lbc_mse(self._p(y), self._p(tgt), self._p(lo), self._p(hi),
                                           self._p(wch_d), self._p(dyb), self._p(loss_part),
                                           nblk, B, P, COUT, norm, self._stream()))
            # layer 2 weights, then back through layer 2, then layer 1 weights
            nat.check(self.lib.fds_lbc_conv_bwd_weight(self._p(dyb), self._p(hb), self._p(part2),
                                                       B, P, HIDDEN, COUT, TAPS, self._stream()))
            nat.check(self.lib.fds_lbc_reduce_weight(self._p(part2), self._p(side), self._p(g2),
                                                     B, 4, COUT * TAPS * HIDDEN, self._stream()))
            nat.check(self.lib.fds_lbc_conv_bwd_data(self._p(dyb), self._p(self.w2t),
                                                     self._p(side), self._p(dh), self._p(dhb),
                                                     B, P, HIDDEN, COUT, TAPS, self._stream()))
            nat.check(self.lib.fds_lbc_conv_bwd_weight(self._p(dhb), self._p(x), self._p(part1),
                                                       B, P, CIN, HIDDEN, TAPS, self._stream()))
            nat.check(self.lib.fds_lbc_reduce_weight(self._p(part1), self._p(side), self._p(g1),
                                                     B, 4, HIDDEN * TAPS * CIN, self._stream()))
            self.adam_step += 1
            for w, g, m1, m2 in ((self.w1, g1, self.opt[0], self.opt[1]),
                                 (self.w2, g2, self.opt[2], self.opt[3])):
                nat.check(self.lib.fds_lbc_adam(self._p(w), self._p(m1), self._p(m2), self._p(g),
                                                w.numel(), lr, 0.9, 0.999, 1e-8, self.adam_step,
                                                self._stream()))
            self._pack()
            if log_every and (it % log_every == 0 or it == iters - 1):
                history.append((it, float(loss_part.sum().item())))
        self.torch.cuda.synchronize(self.dev)
        return history

    # ------------------------------------------------------------------ inference
    def attach(self, fld):
        """Switch the learned layer on for a G x G FP32 Acoustic2D field."""
        if not fld.matrices_assembled:
            fld.assemble_matrices()
        runner = fld._runner()
        runner.sync(fld)
        self.torch.cuda.synchronize(self.dev)
        cs = self.cs.ctypes.data_as(ct.c_void_p)
        nat.check(self.lib.fds_set_learned_boundary(runner.ctx, self._p(self.w1b),
                                                    self._p(self.w2b), LAYER, cs))
        return fld

    # ------------------------------------------------------------------ judging
    def evaluate(self, scene, after=1200, chunk=25, window_margin=0, timing=True):
        """Run the enlarged plain grid, the learned grid and the same grid without the net in
        lock step; reflected energy of both small grids over the interior from the wave's
        arrival at the layer to ``after`` steps later; wall-clock of both methods (unless
        ``timing`` is False)."""
        t_hit = self.wave_arrival(scene, window_margin)
        steps = t_hit + after
        big = self.field(scene, self.big, self.off, steps)
        lrn = self.attach(self.field(scene, self.G, 0, steps))
        abl = self.field(scene, self.G, 0, steps)
        for f in (big, abl):
            f.assemble_matrices()
        acc = {'learned': [0.0, 0.0], 'plain_small': [0.0, 0.0]}
        while big.step < steps:
            n = min(chunk, steps - big.step)
            for f in (big, lrn, abl):
                f.simulate(n)
            if big.step >= t_hit:
                for key, f in (('learned', lrn), ('plain_small', abl)):
                    num, den = ct.c_double(), ct.c_double()
                    nat.check(self.lib.fds_window_sq_diff(
                        f._runner().ctx, 0, big._runner().ctx, 0, self.n_int, self.n_int,
                        LAYER, LAYER, self.off + LAYER, self.off + LAYER, ct.byref(num),
                        ct.byref(den)))
                    acc[key][0] += num.value
                    acc[key][1] += den.value
        res = {k: v[0] / v[1] for k, v in acc.items()}
        res['window'] = (t_hit, steps)
        res['finite'] = bool(np.isfinite(np.asarray(lrn.pressure.values)).all())
        for f in (big, lrn, abl):
            f._runner().close()
        if timing:
            res.update(self.time_methods(scene, steps))
        return res

    def time_methods(self, scene, steps):
        """Wall-clock of the plain enlarged run and of the learned-boundary run (setup and
        shape tuning excluded, both FP32 with HIP-graph chunks)."""
        torch = self.torch
        out = {}
        for key, make in (('enlarged', lambda: self.field(scene, self.big, self.off, steps)),
                          ('learned', lambda: self.attach(self.field(scene, self.G, 0, steps)))):
            f = make()
            f.assemble_matrices()
            # each method on its own best kernel shape (tuning restores the field's state)
            autotune.tune_field(f, steps=20, use_cache=False)
            f.simulate(min(20, steps))           # warm-up (graph capture, first launches)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f.simulate(steps - f.step)
            torch.cuda.synchronize()
            out[key + '_s'] = time.perf_counter() - t0
            out[key + '_steps'] = steps - 20
            f._runner().close()
        out['speedup'] = out['enlarged_s'] / out['learned_s']
        return out

    def validate(self, scene, tolerance=0.01, **kw):
        res = self.evaluate(scene, **kw)
        self.validated = res['learned'] if res['finite'] else float('inf')
        self.tolerance = tolerance
        return res

    def run(self, scene, steps):
        """Interior pressure after ``steps`` steps: learned grid if the validated reflected
        energy is within tolerance, otherwise the plain enlarged-domain run (fallback)."""
        use_net = self.validated is not None and self.validated <= getattr(self, 'tolerance',
                                                                           0.01)
        if use_net:
            f = self.attach(self.field(scene, self.G, 0, steps))
            f.simulate(steps)
            p = np.asarray(f.pressure.values).reshape(self.G, self.G)
            inner = p[LAYER:LAYER + self.n_int, LAYER:LAYER + self.n_int]
        else:
            f = self.field(scene, self.big, self.off, steps)
            f.simulate(steps)
            p = np.asarray(f.pressure.values).reshape(self.big, self.big)
            o = self.off + LAYER
            inner = p[o:o + self.n_int, o:o + self.n_int]
        f._runner().close()
        return inner.copy(), ('learned' if use_net else 'enlarged')
