"""Ideal OFDM channel state information, restated in float64.

``reference_ideal_csi`` follows ``OFDMIdealChannelEstimation.estimate_channel``
(``hermespy/simulation/modem/channel_estimation.py:252-295``) and
``__extract_channel`` (:215-250): the dense impulse response of
``MultipathFadingSample.state(M, M)`` (``fading.py:345-369``), prefix removal,
the mean over each word's N samples, and an un-normalised FFT over the taps.

``closed_form_ideal_csi`` is the same quantity without the dense tensor:
the per-word mean of each sinusoid is a Dirichlet kernel,
mean_{m=m0}^{m0+N-1} e^{j(theta m + phi)} = e^{j(theta (m0 + (N-1)/2) + phi)} sin(N theta/2) / (N sin(theta/2)).
"""

from __future__ import annotations

import numpy as np

from . import fading, ofdm


def reference_ideal_csi(sample: fading.FadingSample, frame: ofdm.Frame) -> np.ndarray:
    """(Nrx, Ntx, words, N) ideal CSI the reference's equaliser sees."""

    m = frame.samples_per_frame
    dense = fading.state(sample, m, m)  # (Nrx, Ntx, M, T)
    picked = ofdm.pick_samples(frame, dense.transpose((0, 1, 3, 2)))  # (Nrx, Ntx, T, words, N)
    avg = picked.mean(axis=-1)  # (Nrx, Ntx, T, words)
    return ofdm.backward_transformation(frame, avg.transpose((0, 1, 3, 2)), normalize=False)


def dirichlet_mean(theta: np.ndarray, n: int) -> np.ndarray:
    """sin(n theta / 2) / (n sin(theta / 2)), equal to 1 at theta = 0 (mod 2 pi)."""

    theta = np.asarray(theta, dtype=np.float64)
    den = n * np.sin(0.5 * theta)
    out = np.ones_like(theta)
    nz = np.abs(den) > 0
    out[nz] = np.sin(0.5 * n * theta[nz]) / den[nz]
    return out


def word_mean_taps(sample: fading.FadingSample, frame: ofdm.Frame) -> np.ndarray:
    """Per-path, per-word mean of the tap over the word's N data samples, (L, words)."""

    p = sample.params
    n = frame.num_subcarriers
    starts = frame.word_starts() + frame.prefix  # first data sample of each word
    centre = starts + 0.5 * (n - 1)
    k = 1 + np.arange(p.num_sinusoids)
    lg, ng = p.los_gains(), p.nlos_gains()
    out = np.zeros((p.num_paths, frame.words), dtype=np.complex128)
    for l in range(p.num_paths):
        # Phase per sample of each sinusoid: omega * cos(.) / fs  (fading.py:303,330)
        th = p.doppler * np.cos((2 * np.pi * k + sample.nlos_angles[l]) / p.num_sinusoids) / sample.bandwidth
        th_los = p.los_doppler * np.cos(sample.los_angles[l]) / sample.bandwidth
        nl = ng[l] * np.sum(np.exp(1j * (np.outer(centre, th) + sample.nlos_phases[l][None, :])) * dirichlet_mean(th, n)[None, :], axis=1)
        lo = lg[l] * np.exp(1j * (centre * th_los + sample.los_phases[l])) * dirichlet_mean(np.array([th_los]), n)[0]
        out[l] = (nl + lo) * (p.gain * p.power_profile[l]) ** 0.5
    return out


def closed_form_ideal_csi(sample: fading.FadingSample, frame: ofdm.Frame) -> np.ndarray:
    """(Nrx, Ntx, words, N) ideal CSI from the Dirichlet closed form."""

    n = frame.num_subcarriers
    m = frame.samples_per_frame
    num_taps = min(1 + round(float(sample.params.delays.max()) * sample.bandwidth), m)
    abar = word_mean_taps(sample, frame)  # (L, words)
    d = fading.delay_samples(sample)
    keep = (d < min(num_taps, n))  # taps beyond the FFT length are truncated by fft(n=N)
    bins = ofdm.subcarrier_bins(n, frame.dc_suppression)
    tw = np.exp(-2j * np.pi * np.outer(d[keep], bins) / n)  # (L', N)
    c = abar[keep].T @ tw  # (words, N)
    return sample.spatial_response[:, :, None, None] * c[None, None, :, :]
