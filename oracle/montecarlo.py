"""Monte Carlo grid-section control, sequential and in float64: the confidence-based
early stopping of ``ScalarEvaluationResult`` and the per-section sample loop.

Test infrastructure: only ``tests/`` use this module, to check
``hermespy_amd.campaign`` against it.

Restated from:

* ``hermespy/core/pymonte/scalar.py:100-125`` (``add_artifact``): running sum, squared
  sum and count per section; every ``min_num_samples`` artifacts the sample standard
  deviation ``std`` gives ``confidence = 2 (1 - Phi(sqrt(count) accuracy / std))`` and
  the section is confident when ``confidence < threshold`` (Bayer et al., Algorithm 1);
  ``std == 0`` never stops.  ``Phi`` is ``scipy.stats.norm.cdf`` (``scalar.py:127-137``;
  ``BitErrorEvaluator._scalar_cdf`` at ``modem/evaluators.py:268-270`` is never called by
  the result).
* ``hermespy/core/pymonte/scalar.py:56-80``: the argument checks.
* ``hermespy/core/pymonte/actors.py:77-108`` (``next_batch``): sections are served round
  robin in grid order; a section leaves the queue once ``num_samples`` tasks were issued,
  and ``actors.py:216-225`` removes it once it is confident.
* ``BitErrorEvaluation.artifact`` (``modem/evaluators.py:231-233``): the artifact of a drop
  is its bit error rate ``errors / bits``.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.stats import norm


def check_arguments(accuracy: float, confidence: float, min_num_samples: int) -> None:
    if confidence > 1.0 or confidence < 0.0:
        raise ValueError("confidence must be between zero and one")
    if accuracy < 0.0:
        raise ValueError("accuracy must be non-negative")
    if min_num_samples < 2:
        raise ValueError("min_num_samples must be at least two")


def confident(total: float, squared: float, count: int, accuracy: float, threshold: float) -> bool:
    """The stopping test of ``scalar.py:117-123`` on one section's running sums."""

    std = ((squared - (total**2 / count)) / (count - 1)) ** 0.5
    if not std > 0.0:
        return False
    return 2 * (1 - norm.cdf(count**0.5 * accuracy / std)) < threshold


class Section:
    """``ScalarEvaluationResult`` state of one grid section (``scalar.py:90-125``)."""

    def __init__(self, accuracy: float = 0.0, confidence: float = 1.0, min_num_samples: int = 1024) -> None:
        check_arguments(accuracy, confidence, min_num_samples)
        self.accuracy, self.threshold, self.min_num_samples = accuracy, confidence, min_num_samples
        self.total, self.squared, self.count = 0.0, 0.0, 0

    def add_artifact(self, scalar: float, compute_confidence: bool = True) -> bool:
        self.total += scalar
        self.squared += scalar**2
        self.count += 1
        if compute_confidence and self.count % self.min_num_samples == 0:
            return confident(self.total, self.squared, self.count, self.accuracy, self.threshold)
        return False

    def mean(self) -> float:
        return self.total / self.count if self.count else math.nan


def run_sections(scalar_of, num_sections: int, num_samples: int, accuracy: float = 0.0, confidence: float = 1.0,
                 min_num_samples: int = 1024) -> tuple:
    """Each section draws samples k = 0, 1, ... (``scalar_of(section, k)``) until it is
    confident or has ``num_samples``; returns (counts, means, sums, squared sums)."""

    counts = np.zeros(num_sections, dtype=np.int64)
    means = np.zeros(num_sections)
    sums = np.zeros(num_sections)
    squares = np.zeros(num_sections)
    for s in range(num_sections):
        sec = Section(accuracy, confidence, min_num_samples)
        for k in range(num_samples):
            if sec.add_artifact(float(scalar_of(s, k))):
                break
        counts[s], means[s], sums[s], squares[s] = sec.count, sec.mean(), sec.total, sec.squared
    return counts, means, sums, squares
