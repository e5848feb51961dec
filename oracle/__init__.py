"""CPU oracle for the Monte Carlo link-drop path (TEST INFRASTRUCTURE ONLY).

This package restates, in NumPy/SciPy float64, the algorithm of the reference
drop path: SoS multipath fading (``hermespy/channel/fading/fading.py``), OFDM
(``hermespy/modem/waveforms/orthogonal/{ofdm,waveform}.py``), 3GPP Gray
mapping/demapping (``hermespy/modem/tools/psk_qam_mapping.py``), ideal CSI
(``hermespy/simulation/modem/channel_estimation.py``), ZF equalisation
(``hermespy/modem/waveform.py``), AWGN (``hermespy/simulation/rf/noise``) and the
bit error evaluator (``hermespy/modem/evaluators.py``).  Each function cites the
reference file:line it follows.

It checks the HIP path and is never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``hermespy_amd``) never calls into it.

Pinning (what checks each part independently of this restatement):

* ``philox``     -- Random123 Philox4x32-10 known-answer vectors.
* ``mapping``    -- the reference's own PSK/QAM known-answer vectors
                    (``tests/unit_tests/modem/tools/test_psk_qam_mapping.py:80-144``),
                    exhaustive demap brute force against nearest-point decisions.
* ``ofdm``       -- independent O(N^2) DFT on small N, round-trip and power invariants.
* ``fading``     -- closed forms (K=inf pure LOS phasor, omega=0 constant, gain
                    scaling, delay shift), direct convolution brute force.
* ``csi``        -- dense-CIR average (reference) == Dirichlet closed form.
* ``equalize``   -- ``numpy.linalg.pinv``/``solve`` (library), MMSE -> ZF as sigma->0.
                    MMSE has no reference counterpart: parity against the reference
                    is unpinned; it is pinned by the float64 closed form only.
* ``ber``        -- XOR-popcount == sum(abs(tx - rx)).
"""
