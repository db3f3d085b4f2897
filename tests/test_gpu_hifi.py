"""HiFi Hilbert FM discriminator on the HIP device (four-step FFT kernels +
discriminator, through vhsdecode_hip_hifi_demodulate_host) against the oracle's
restatement of HiFiHilbertDiscriminator.Demodulate, which the reference's float-bit
vectors pin (tests/test_laserdisc_oracle.py).  Block of 2^22 samples at the 2^23 Hz
Hilbert IF rate, as the reference decodes.  Runs on an MI355X only."""
import numpy as np
import pytest

from oracle import oracle as O
from vhsdecode_hip.hifi import HipHiFiHilbertDiscriminator

pytestmark = pytest.mark.gpu

IF_RATE = 1 << 23
BLOCK = 1 << 22


def _fm(n, carrier, deviation, tone, seed, snr_db=40.0):
    t = np.arange(n, dtype=np.float64) / IF_RATE
    inst = carrier + deviation * np.sin(2 * np.pi * tone * t)
    x = np.cos(2 * np.pi * np.cumsum(inst) / IF_RATE)
    rng = np.random.default_rng(seed)
    x += rng.normal(0.0, (1 / np.sqrt(2)) / 10 ** (snr_db / 20), n)
    return x.astype(np.float32)


@pytest.fixture(scope="module", params=[(1_300_000.0, 440.0, 1), (1_700_000.0, 880.0, 2)], ids=["left", "right"])
def case(request):
    carrier, tone, seed = request.param
    x = _fm(BLOCK, carrier, 150_000.0, tone, seed)
    ref = np.full(BLOCK, 123.25, dtype=np.float32)
    O.hifi_demodulate(x, ref, IF_RATE, int(carrier), 150_000)
    got = np.full(BLOCK, 123.25, dtype=np.float32)
    with HipHiFiHilbertDiscriminator(IF_RATE, carrier, 150_000.0, BLOCK) as d:
        d.demodulate(x, got)
    print(f"\n[HiFi {carrier / 1e6:.1f} MHz] block={BLOCK}", flush=True)
    return dict(ref=ref, got=got, carrier=carrier)


def test_last_sample_untouched(case):
    assert case["got"][-1] == np.float32(123.25)


def test_phase_wrap_pattern(case):
    """The output is -carrier/deviation except where a phase step wraps; the wraps must be
    the same samples (one per carrier cycle)."""
    ref, got = case["ref"][:-1], case["got"][:-1]
    base = np.float32(-case["carrier"] / 150_000.0)
    wr, wg = ref != base, got != base
    mismatch = int(np.count_nonzero(wr != wg))
    cycles = case["carrier"] * BLOCK / IF_RATE
    print(f"wraps ref={int(wr.sum())} gpu={int(wg.sum())} (carrier cycles {cycles:.0f}), mismatched {mismatch}")
    assert abs(int(wr.sum()) - cycles) < 0.01 * cycles
    assert mismatch <= 1e-6 * BLOCK


def test_values_within_one_ulp(case):
    """At the wraps the reference forms the correction as a difference of running FP64
    sums; the device uses the step's own correction.  The two agree to a float32 ulp
    (measured on MI355X: identical bits on both carriers)."""
    base = np.float32(-case["carrier"] / 150_000.0)
    same = (case["ref"][:-1] != base) == (case["got"][:-1] != base)
    ref, got = case["ref"][:-1][same].astype(np.float64), case["got"][:-1][same].astype(np.float64)
    ulp = np.spacing(np.abs(case["ref"][:-1][same])).astype(np.float64)
    diff = np.abs(ref - got)
    frac = float(np.count_nonzero(diff) / diff.size)
    print(f"differing samples {frac:.2e}, max {float(np.max(diff / ulp)):.1f} ulp")
    assert np.all(diff <= ulp)
    assert frac <= 1e-3
