"""Learned pilot-to-grid channel estimator (bfloat16 MFMA convolutions) and its trainer.

A 1-D convolutional network over the 64 pilot positions of one (receive,
transmit) stream pair maps the LS values at that stream's pilots (two pilot
words, real and imaginary parts, a stream flag and the SNR: 6 of 8 input channels) to
the channel of every resource element: 64 positions x (words x N / 64
subcarriers x re/im) outputs.  Four circular layers, kernel 5, dilations 1, 4,
16, 1 (the receptive field covers the whole cyclic pilot axis, which a linear
estimator needs to reach LMMSE: the frequency response is periodic and, without
the DC roll, uniformly sampled by a stream's comb pilots), 64 hidden channels,
GELU (tanh form), plus a linear 17-tap skip convolution from the input to the
output: about 146k parameters.  Weights and activations are bfloat16,
products accumulate in float32 on the MFMA units (``hpx_conv1d_forward``);
training runs on the device with hand-written kernels: backward-data as the
forward convolution with transposed, flipped weights, float32 weight
gradients, batch-NMSE loss and Adam, on drops the fused simulator generates
(``hpx_run_batch_frames`` exports of the received grid and the ideal CSI).

It replaces the interpolation step of ``OrthogonalLeastSquaresChannelEstimation``
(``hermespy/modem/waveforms/orthogonal/channel.py:29-66``); the LMMSE estimator
of :mod:`hermespy_amd.lmmse` is its fallback and the baseline its NMSE is measured against.
"""

from __future__ import annotations

import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .configs import LinkConfig

POSITIONS = 64
TAPS = 5
SKIP_TAPS = 17
HIDDEN = 64
CIN = 8  # 2 pilot words x (re, im) + stream flag + SNR, zero-padded to 8
WEIGHTS_PATH = Path(__file__).resolve().parent / "data" / "learned_estimator.npz"


DILATIONS = (1, 4, 16, 1)


def layer_shapes(cfg: LinkConfig) -> list:
    """(cin, cout, gelu) of the four layers (dilations DILATIONS)."""

    cout = 2 * cfg.words * (cfg.num_subcarriers // POSITIONS)
    return [(CIN, HIDDEN, True), (HIDDEN, HIDDEN, True), (HIDDEN, HIDDEN, True), (HIDDEN, cout, False)]


def skip_shape(cfg: LinkConfig) -> tuple:
    """(cin, cout) of the linear skip convolution (SKIP_TAPS taps)."""

    return CIN, layer_shapes(cfg)[-1][1]


def init_params(cfg: LinkConfig, seed: int = 1) -> list:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights [cout][5][cin] and biases of the four layers,
    then the skip convolution, zero-initialised, float32."""

    rng = np.random.default_rng(seed)
    out = []
    for cin, cout, _ in layer_shapes(cfg):
        bound = 1.0 / math.sqrt(cin * TAPS)
        w = rng.uniform(-bound, bound, (cout, TAPS, cin)).astype(np.float32)
        b = rng.uniform(-bound, bound, cout).astype(np.float32)
        out.append((w, b))
    cin, cout = skip_shape(cfg)
    out.append((np.zeros((cout, SKIP_TAPS, cin), np.float32), np.zeros(cout, np.float32)))
    return out


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""

    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


class ConvNetEstimator:
    """The network on the device: float32 master parameters and their bfloat16 copies."""

    def __init__(self, cfg: LinkConfig, params: list | None = None, device: str | torch.device = "cuda") -> None:
        if cfg.frame_mask is None or cfg.num_subcarriers % POSITIONS:
            raise ValueError("the learned estimator needs a pilot frame with a multiple of 64 subcarriers")
        if cfg.dc_suppression:
            raise ValueError("the learned estimator's circular convolutions need a grid without DC roll")
        self.cfg = cfg
        self.device = torch.device(device)
        self.shapes = layer_shapes(cfg)
        params = init_params(cfg) if params is None else params
        dev = self.device
        self.w = [torch.as_tensor(np.ascontiguousarray(w, dtype=np.float32), device=dev) for w, _ in params]
        self.b = [torch.as_tensor(np.ascontiguousarray(b, dtype=np.float32), device=dev) for _, b in params]
        self.w16 = [torch.empty(w.shape, dtype=torch.bfloat16, device=dev) for w in self.w]
        self.wt16 = [torch.empty((w.shape[2], w.shape[1], w.shape[0]), dtype=torch.bfloat16, device=dev) for w in self.w]
        self.refresh()

    @property
    def num_params(self) -> int:
        return int(sum(w.numel() for w in self.w) + sum(b.numel() for b in self.b))

    def refresh(self) -> None:
        """Rewrite the bfloat16 weight copies from the float32 parameters."""

        lib, st = _lib.load(), _lib.stream_handle()
        for w, w16, wt16 in zip(self.w, self.w16, self.wt16):
            _lib.check(lib.hpx_conv1d_weight_copies(_lib.ptr(w), w.shape[0], w.shape[2], w.shape[1], _lib.ptr(w16),
                                                    _lib.ptr(wt16), st))

    # ------------------------------------------------------------ inference
    def net_input(self, ls: torch.Tensor, noise_power: torch.Tensor) -> torch.Tensor:
        """LS values [num_tx][drops][num_rx][pilot_words * 64] and the drops' noise powers -> bf16 input
        [S][64][8]."""

        num_tx, drops, num_rx, p = ls.shape
        words = p // POSITIONS
        x = torch.empty((num_tx * drops * num_rx, POSITIONS, CIN), dtype=torch.bfloat16, device=self.device)
        _lib.check(_lib.load().hpx_net_input(_lib.ptr(ls), drops, num_rx, num_tx, words, CIN, _lib.ptr(noise_power),
                                             _lib.ptr(x), _lib.stream_handle()))
        return x

    def forward(self, x: torch.Tensor, keep: bool = False):
        """Network output [S][64][cout] (float32); with keep, also the per-layer (input, pre-activation)."""

        lib, st = _lib.load(), _lib.stream_handle()
        s = x.shape[0]
        cache = []
        a = x
        out = None
        for i, (cin, cout, act) in enumerate(self.shapes):
            last = i == len(self.shapes) - 1
            pre = torch.empty((s, POSITIONS, cout), dtype=torch.float32, device=self.device) if (keep or last) else None
            nxt = None if last else torch.empty((s, POSITIONS, cout), dtype=torch.bfloat16, device=self.device)
            _lib.check(lib.hpx_conv1d_forward(_lib.ptr(a), s, cin, TAPS, DILATIONS[i], _lib.ptr(self.w16[i]),
                                              _lib.ptr(self.b[i]), cout, 1 if act else 0, _lib.ptr(pre), _lib.ptr(nxt),
                                              st))
            cache.append((a, pre))
            a = nxt
            out = pre
        # linear skip path from the input, accumulated onto the last layer's output
        cin, cout = skip_shape(self.cfg)
        _lib.check(lib.hpx_conv1d_forward(_lib.ptr(x), s, cin, SKIP_TAPS, 1, _lib.ptr(self.w16[-1]),
                                          _lib.ptr(self.b[-1]), cout, 2, _lib.ptr(out), None, st))
        return (out, cache) if keep else out

    def estimate_grid(self, grid: torch.Tensor, pilot_index: torch.Tensor, pilot_values: torch.Tensor,
                      noise_power: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Channel estimates [num_tx][drops][num_rx][words * N] of received grids [drops][num_rx][words][N]
        in one launch (``hpx_learned_estimate``: pilot LS, the network and the skip path with the
        activations in LDS)."""

        import ctypes

        c = self.cfg
        drops = grid.shape[0]
        if out is None:
            out = torch.empty((c.num_tx, drops, c.num_rx, c.words * c.num_subcarriers), dtype=torch.complex64,
                              device=self.device)
        ws = (ctypes.c_void_p * 5)(*[w.data_ptr() for w in self.w16])
        bs = (ctypes.c_void_p * 5)(*[b.data_ptr() for b in self.b])
        _lib.check(_lib.load().hpx_learned_estimate(
            _lib.ptr(grid), drops, c.num_rx, c.num_tx, c.words, c.num_subcarriers, _lib.ptr(pilot_index),
            _lib.ptr(pilot_values), pilot_index.shape[1], _lib.ptr(noise_power), ctypes.addressof(ws), ctypes.addressof(bs),
            self.shapes[-1][1], _lib.ptr(out), _lib.stream_handle()))
        return out

    def estimate(self, ls: torch.Tensor, noise_power: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Channel estimates [num_tx][drops][num_rx][words * N] from LS values (layer by layer, the
        training path's arithmetic)."""

        c = self.cfg
        num_tx, drops, num_rx, _ = ls.shape
        y = self.forward(self.net_input(ls, noise_power))
        if out is None:
            out = torch.empty((num_tx, drops, num_rx, c.words * c.num_subcarriers), dtype=torch.complex64,
                              device=self.device)
        _lib.check(_lib.load().hpx_net_output(_lib.ptr(y), y.shape[0], c.words, c.num_subcarriers, _lib.ptr(out),
                                              _lib.stream_handle()))
        return out

    # ------------------------------------------------------------ persistence
    def save(self, path: str | Path) -> None:
        arrays = {}
        for i, (w, b) in enumerate(zip(self.w, self.b)):
            arrays[f"w{i}"] = w.cpu().numpy()
            arrays[f"b{i}"] = b.cpu().numpy()
        np.savez(path, **arrays)

    @classmethod
    def load(cls, cfg: LinkConfig, path: str | Path = WEIGHTS_PATH, device: str | torch.device = "cuda"):
        with np.load(path, allow_pickle=False) as f:
            params = [(f[f"w{i}"], f[f"b{i}"]) for i in range(5)]
        if params[-1][0].shape[1] != SKIP_TAPS:
            raise ValueError(f"{path}: skip convolution with {params[-1][0].shape[1]} taps, expected {SKIP_TAPS}")
        return cls(cfg, params, device)


class Trainer:
    """Adam on the batch NMSE of drops generated by the fused simulator."""

    def __init__(self, net: ConvNetEstimator, drops_per_step: int = 1024, lr: float = 2e-3, seed: int = 7,
                 betas=(0.9, 0.999), eps: float = 1e-8) -> None:
        from .simulation import BatchedLink

        self.net = net
        self.cfg = net.cfg
        self.link = BatchedLink(self.cfg, net.device)
        from .lmmse import stream_pilots

        idx, val = stream_pilots(self.cfg)
        self.pilot_index = torch.as_tensor(idx, device=net.device)
        self.pilot_values = torch.as_tensor(val.astype(np.complex64), device=net.device)
        self.num_pilots = idx.shape[1]
        self.drops = drops_per_step
        self.lr = lr
        self.seed = seed
        self.betas = betas
        self.eps = eps
        self.step_count = 0
        dev = net.device
        self.m1 = [torch.zeros_like(p) for p in net.w + net.b]
        self.m2 = [torch.zeros_like(p) for p in net.w + net.b]
        self.gw = [torch.empty_like(w) for w in net.w]
        self.gb = [torch.empty_like(b) for b in net.b]
        s = self.cfg.num_tx * drops_per_step * self.cfg.num_rx
        lib = _lib.load()
        ws = max([int(lib.hpx_conv1d_wgrad_workspace(s, cin, cout, TAPS)) for cin, cout, _ in net.shapes] +
                 [int(lib.hpx_conv1d_wgrad_workspace(s, *skip_shape(self.cfg), SKIP_TAPS))])
        self.workspace = torch.empty(ws // 4 + 1, dtype=torch.float32, device=dev)
        c = self.cfg
        self.grid = torch.empty((drops_per_step, c.num_rx, c.words, c.num_subcarriers), dtype=torch.complex64,
                                device=dev)
        self.csi = torch.empty((drops_per_step, c.words, c.num_subcarriers), dtype=torch.complex64, device=dev)
        self.spatial = torch.empty((drops_per_step, c.num_rx, c.num_tx), dtype=torch.complex64, device=dev)
        self.errors = torch.empty(drops_per_step, dtype=torch.int32, device=dev)
        self.num_groups = len(self.cfg.snr_db)
        self.work = torch.zeros(3 * self.num_groups, dtype=torch.float64, device=dev)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)

    def simulate(self, first_drop: int):
        """Generate drops first_drop.. (SNR point i % len(snr_db), Doppler per drop): LS values and the truth."""

        c = self.cfg
        n = self.drops
        snr = np.asarray(c.snr_db)
        npw = torch.as_tensor(c.noise_power(snr[(first_drop + np.arange(n)) % snr.size]), dtype=torch.float64,
                              device=self.net.device)
        self.noise_power = npw
        self.link.run(self.seed, first_drop, npw, self.errors, doppler=c.draw_doppler(first_drop, n), grid=self.grid,
                      csi=self.csi, spatial=self.spatial)
        ls = torch.empty((c.num_tx, n, c.num_rx, self.num_pilots), dtype=torch.complex64, device=self.net.device)
        _lib.check(_lib.load().hpx_pilot_ls(_lib.ptr(self.grid), n, c.num_rx, c.num_tx, c.words * c.num_subcarriers,
                                            self.num_pilots, _lib.ptr(self.pilot_index), _lib.ptr(self.pilot_values),
                                            _lib.ptr(ls), _lib.stream_handle()))
        return ls

    def step(self, first_drop: int, lr: float | None = None) -> torch.Tensor:
        """One Adam step on drops first_drop .. first_drop + drops_per_step; returns the loss, the mean
        over SNR points of the NMSE in dB (device scalar)."""

        net, c = self.net, self.cfg
        lib, st = _lib.load(), _lib.stream_handle()
        ls = self.simulate(first_drop)
        x = net.net_input(ls, self.noise_power)
        out, cache = net.forward(x, keep=True)
        grad = torch.empty_like(out)
        group = torch.as_tensor(((first_drop + np.arange(self.drops)) % self.num_groups).astype(np.int32),
                                device=net.device)
        self.work.zero_()
        _lib.check(lib.hpx_nmse_grad(_lib.ptr(out), _lib.ptr(self.csi), _lib.ptr(self.spatial), self.drops, c.num_rx,
                                     c.num_tx, c.words, c.num_subcarriers, _lib.ptr(group), self.num_groups,
                                     _lib.ptr(self.work), _lib.ptr(grad), _lib.ptr(self.loss), st))
        s = x.shape[0]
        for i in range(len(net.shapes) - 1, -1, -1):
            cin, cout, act = net.shapes[i]
            a_in, pre = cache[i]
            dz = torch.empty((s, POSITIONS, cout), dtype=torch.bfloat16, device=net.device)
            if act:
                _lib.check(lib.hpx_gelu_backward(_lib.ptr(grad), _lib.ptr(pre), grad.numel(), _lib.ptr(dz), st))
            else:
                _lib.check(lib.hpx_to_bf16(_lib.ptr(grad), grad.numel(), _lib.ptr(dz), st))
            _lib.check(lib.hpx_conv1d_wgrad(_lib.ptr(dz), _lib.ptr(a_in), s, cin, cout, TAPS, DILATIONS[i],
                                            _lib.ptr(self.workspace), _lib.ptr(self.gw[i]), _lib.ptr(self.gb[i]), st))
            if i == len(net.shapes) - 1:  # the skip convolution sees the same output gradient
                scin, scout = skip_shape(c)
                _lib.check(lib.hpx_conv1d_wgrad(_lib.ptr(dz), _lib.ptr(x), s, scin, scout, SKIP_TAPS, 1,
                                                _lib.ptr(self.workspace), _lib.ptr(self.gw[-1]), _lib.ptr(self.gb[-1]),
                                                st))
            if i > 0:
                grad = torch.empty((s, POSITIONS, cin), dtype=torch.float32, device=net.device)
                _lib.check(lib.hpx_conv1d_forward(_lib.ptr(dz), s, cout, TAPS, DILATIONS[i], _lib.ptr(net.wt16[i]), None,
                                                  cin, 0, _lib.ptr(grad), None, st))
        self.step_count += 1
        b1, b2 = self.betas
        for p, g, m1, m2 in zip(net.w + net.b, self.gw + self.gb, self.m1, self.m2):
            _lib.check(lib.hpx_adam_step(_lib.ptr(p), _lib.ptr(g), _lib.ptr(m1), _lib.ptr(m2), p.numel(),
                                         float(self.lr if lr is None else lr), b1, b2, self.eps, self.step_count, st))
        net.refresh()
        return self.loss[0]


def train(cfg: LinkConfig, num_drops: int, drops_per_step: int = 1024, lr: float = 2e-3, log_every: int = 50,
          first_drop: int = 1 << 30, params: list | None = None, verbose: bool = True,
          save_to: str | Path | None = None, save_every: int = 100) -> ConvNetEstimator:
    """Train on num_drops simulator drops (drop indices from first_drop, disjoint from evaluation drops);
    cosine learning-rate decay."""

    net = ConvNetEstimator(cfg, params)
    tr = Trainer(net, drops_per_step, lr)
    steps = max(1, num_drops // drops_per_step)
    t0 = time.perf_counter()
    for k in range(steps):
        rate = lr * 0.5 * (1 + math.cos(math.pi * k / steps))
        loss = tr.step(first_drop + k * drops_per_step, rate)
        if verbose and (k % log_every == 0 or k == steps - 1):
            print(f"step {k + 1}/{steps} mean nmse {float(loss):.2f} dB "
                  f"lr {rate:.2e} {time.perf_counter() - t0:.1f} s", flush=True)
        if save_to is not None and (k + 1) % save_every == 0:
            net.save(save_to)
    return net


def _main() -> None:
    import argparse

    from .configs import mimo2x2_512_pilots

    ap = argparse.ArgumentParser(description="train the learned channel estimator on simulator drops")
    ap.add_argument("--drops", type=int, default=1_000_000)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", type=str, default=str(WEIGHTS_PATH))
    args = ap.parse_args()
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    net = train(mimo2x2_512_pilots(), args.drops, args.batch, args.lr, save_to=args.out)
    torch.cuda.synchronize()
    net.save(args.out)
    print(f"saved {net.num_params} parameters to {args.out}", flush=True)


if __name__ == "__main__":
    _main()
