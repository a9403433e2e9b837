"""Raw echoes to an encoded RSLC block with every step on the GPU and the
data resident in HBM between steps: presum (gap filling onto the uniform
pulse grid) -> range compression -> split-precision backprojection -> RSLC
block finishing (range correction, statistics, complex32), against the
chained oracles (oracle/presum.py, oracle/rangecomp.py, the C++ backprojection
oracle in FP64 mode, oracle/rslc.py).

Bar: relative error <= 1e-3 and RMS phase <= 1e-3 rad on the focused image
(the split-precision bar; the steps before it agree to ~1e-6), complex32
encodings within one half-precision rounding of the oracle chain's."""
import numpy as np
import pytest

import oracle
import presum as po
import rangecomp as rco
import rslc as ro
from helpers import complex_rel_error, phase_errors, rms, scene_and_raw

pytestmark = pytest.mark.gpu


def test_raw_to_rslc_on_the_gpu(torch_gpu):
    torch = torch_gpu
    from isce3_amd.focus import RangeComp, backproject, form_linear_chirp
    from isce3_amd.presum import resample
    from isce3_amd.rslc import BlockFinisher
    from isce3_amd.sim import SPEED
    sc, rc_true = scene_and_raw("tiny")
    nt, nr = rc_true.shape
    grid = sc.in_grid
    # echoes: the range-compressed scene spread by a short chirp, with a
    # 12-bin blanking gap that walks across the swath pulse by pulse
    chirp = form_linear_chirp(3e12, 2.0e-6, 24e6)          # 49 samples
    m = len(chirp)
    echo = np.stack([np.convolve(rc_true[k], chirp) for k in range(nt)]).astype(np.complex64)
    ne = echo.shape[1]
    t = grid.sensing_start + np.arange(nt) / grid.prf
    g = 30 + (np.arange(nt) * 41) % (ne - 60)
    sw = np.zeros((2, nt, 2), np.int64)
    sw[0, :, 1] = g
    sw[1, :, 0], sw[1, :, 1] = g + 12, ne
    echo_grid = type(grid)(grid.sensing_start, grid.wavelength, grid.prf, grid.starting_range,
                           grid.range_pixel_spacing, grid.lookside, nt, ne, grid.ref_epoch)
    print(f"[pipeline] echoes {echo.shape}, chirp {m}", flush=True)

    # --- GPU chain, device tensors throughout
    d_echo = torch.from_numpy(echo).cuda()
    d_uniform = resample(d_echo, t, echo_grid, sw, sc.orbit, sc.doppler)
    comp = RangeComp(chirp, ne, maxbatch=nt, mode=RangeComp.Mode.Full)
    d_full = torch.empty((nt, comp.output_size), dtype=torch.complex64, device="cuda:0")
    comp.rangecompress(d_full, d_uniform)
    d_rc = d_full[:, m - 1:m - 1 + nr].contiguous()           # the matched filter's peak lag
    d_img = torch.empty(sc.out_grid.shape, dtype=torch.complex64, device="cuda:0")
    failed = backproject(d_img, sc.out_geometry(), d_rc, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                         sc.kernel, sc.dry_tropo_model, precision="split")
    assert not failed
    # range correction: a phase ramp scaled so the image fits half precision
    scale = 1.0 / float(torch.abs(d_img).max())
    ramp = (scale * np.exp(0.3j * np.arange(sc.out_grid.width))).astype(np.complex64)
    d_block = d_img.clone()
    enc, st = BlockFinisher(d_block.numel()).finish(d_block, range_cor=torch.from_numpy(ramp).cuda(),
                                                   data_type="complex32")

    # --- oracle chain
    uni = po.resample(echo, t, grid.prf, grid.starting_range, grid.range_pixel_spacing, sw,
                      lambda tt: SPEED,                 # the straight orbit's constant speed
                      lambda tt, r: np.zeros_like(r))
    assert complex_rel_error(d_uniform.cpu().numpy(), uni) < 1e-5
    full = rco.rangecompress(chirp, uni.astype(np.complex64), rco.FULL, "fp64")
    rc = full[:, m - 1:m - 1 + nr].astype(np.complex64)
    assert complex_rel_error(d_rc.cpu().numpy(), rc) < 1e-5
    img, _, _, rcode = oracle.backproject(sc.out_geometry(), rc, sc.in_geometry(), sc.dem, sc.fc,
                                          sc.ds, sc.kernel, sc.dry_tropo_model, mode=1,
                                          kernel_bandwidth=0.8)
    assert rcode == 0
    z = d_img.cpu().numpy()
    rel, ph = complex_rel_error(z, img), rms(phase_errors(z, img))
    print(f"[pipeline] image rel={rel:.2e} rms_phase={ph:.2e}", flush=True)
    assert rel <= 1e-3 and ph <= 1e-3
    # the finisher on the GPU image equals the oracle finisher on the same image
    zo, enc_o, ex = ro.finish_block(z, ramp, None, "complex32")
    got = enc.view(torch.int32).cpu().numpy().view(np.uint16)
    assert np.array_equal(got, enc_o.view(np.uint16))
    assert st.n_valid == ex["n_valid"] == z.size
    # and the encoded image is the oracle image to complex32 precision
    dec = (enc_o["r"].astype(np.float32) + 1j * enc_o["i"].astype(np.float32))
    ref_scaled = img * ramp[None, :]
    assert complex_rel_error(dec, ref_scaled) <= 2e-3
