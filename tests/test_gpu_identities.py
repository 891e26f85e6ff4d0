"""The two hardware identities behind the projectors' exact fast paths, over
every float32 bit pattern on the device (xpr_check_fast_paths):

* v_rcp_f32 plus one Newton FMA equals the IEEE quotient 1/w for every normal
  |w| in [1e-6, 1e6] (outside that range the kernels divide);
* v_cvt_rpi_i32_f32 (floor(u + 0.5)) equals (int)roundf(u) for every u in
  (-0.5, 2^23), the coordinates the canonical bounds test can accept.

The mismatches outside the reciprocal's range must exist (the check would
otherwise not look at the inputs it claims to)."""
import pytest

pytestmark = pytest.mark.gpu


def test_fast_path_identities_hold_for_every_float(cuda):
    import xpulse_recon_amd as X
    rcp_in, rpi, rcp_out = X.check_fast_paths()
    assert rcp_in == 0
    assert rpi == 0
    assert rcp_out > 0
