"""GPU parity of the FP64 tracer and the post-processing against the FP64
CPU oracle, through the C ABI (libkerrtrace.so)."""
import sys

import numpy as np
import pytest

from kerrtrace_amd import config as K
from oracle import oracle as O
from tests.parity import compare, ldr_diff

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _render_gpu(torch, cfg, cam):
    from kerrtrace_amd.renderer import HipRenderer
    with HipRenderer(cfg) as r:
        hdr, cls = r.trace(cam)
        torch.cuda.synchronize()
        st = r.stats()
        return hdr.cpu().numpy(), cls.cpu().numpy(), st


# Tolerances of the FP64 mode: hit classes may differ only on boundary
# pixels and on at most 0.1% of pixels; HDR relative error <= 1e-5 on
# pixels of equal class (the HDR buffer itself is float32, 6e-8).
FP64_MAX_MISMATCH_FRAC = 1e-3
FP64_MAX_REL = 1e-5


@pytest.mark.parametrize("name,size", [("cpu256_a09", None), ("hd1080_a0998", (480, 270))])
def test_fp64_frame_matches_oracle(torch_gpu, name, size):
    """The 256x256 workload at its full size and the 1080p scene at 480x270:
    within the parity bar, and in fact bit for bit the oracle's frame (every
    hit class and HDR float, every work counter)."""
    cfg, cam = K.preset(name)
    if size is not None:
        cfg, cam = K.scaled(cfg, cam, *size)
    say(f"[parity] fp64 {name} {cfg.width}x{cfg.height}")
    g_hdr, g_cls, st = _render_gpu(torch_gpu, cfg, cam)
    o_hdr, o_cls, o_att, ost = O.render(cfg, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[parity] {name}: {m}; gpu attempts {st['attempts']} oracle {ost.attempts}")
    assert m["class_mismatch_non_boundary"] == 0
    assert m["class_mismatch_frac"] <= FP64_MAX_MISMATCH_FRAC
    assert m["max_rel"] <= FP64_MAX_REL
    assert np.array_equal(g_cls, o_cls)
    assert np.array_equal(g_hdr[..., :3], o_hdr.astype(np.float32))
    assert int(st["attempts"]) == int(ost.attempts) and int(st["crossings"]) == int(ost.crossings)
    assert np.all(g_hdr[..., 3] == 1.0)


@pytest.mark.parametrize("spin,noise,spp", [(-0.7, 1, 1), (0.0, 0, 2), (0.999, 1, 1)])
def test_fp64_scene_variants_match_oracle_bitwise(torch_gpu, spin, noise, spp):
    """A retrograde spin (disk from the retrograde ISCO, negative frame
    term), a Schwarzschild hole with the noise off and two samples per
    pixel, and a near-extremal spin: every hit class and HDR float equal to
    the oracle's."""
    cfg, cam = K.scaled(*K.preset("cpu256_a09"), 96, 80)
    cfg.spin = spin
    cfg.noise_enabled = noise
    cfg.spp = spp
    say(f"[parity] variant spin={spin} noise={noise} spp={spp}")
    g_hdr, g_cls, st = _render_gpu(torch_gpu, cfg, cam)
    o_hdr, o_cls, _, ost = O.render(cfg, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[parity] variant: {m}; attempts gpu {st['attempts']} oracle {ost.attempts}")
    assert np.array_equal(g_cls, o_cls)
    # every HDR float equal to the oracle's FP64 value rounded to FP32
    assert np.array_equal(g_hdr[..., :3], o_hdr.astype(np.float32))
    assert int(st["attempts"]) == int(ost.attempts)


@pytest.mark.parametrize("name,spin", [("hd1080_a0998", None), ("uhd2160_i60", 0.0), ("uhd2160_i60", 0.5),
                                       ("uhd2160_i60", 0.9), ("uhd2160_i60", 0.998)])
def test_fp64_full_size_frame_matches_oracle(torch_gpu, name, spin):
    """The measured frames at their full size (1920x1080 at tol 1e-9;
    3840x2160 at i = 60 deg, tol 1e-7, in FP64, at the four spins the mixed
    tracer is compared at): the oracle finishes them in seconds on the host's
    threads.  The frames are the oracle's bit for bit: every hit class, every
    attempt count and every HDR float."""
    cfg, cam = K.preset(name) if spin is None else K.preset(name, spin=spin)
    cfg.precision = K.KT_PRECISION_FP64
    say(f"[parity] fp64 full size {name} {cfg.width}x{cfg.height}")
    g_hdr, g_cls, st = _render_gpu(torch_gpu, cfg, cam)
    o_hdr, o_cls, _, ost = O.render(cfg, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[parity] {name} full: {m}; gpu attempts {st['attempts']} oracle {ost.attempts} "
        f"({ost.seconds:.1f} s on {ost.threads} threads)")
    assert m["class_mismatch"] == 0 and m["subclass_mismatch"] == 0
    assert m["exact_fp32_frac"] == 1.0
    assert st["attempts"] == ost.attempts and st["crossings"] == ost.crossings
    assert st["rhs_evals"] == ost.rhs_evals and st["accepted"] == ost.accepted
    assert np.all(g_hdr[..., 3] == 1.0)


def test_secant_reference_root_mode(torch_gpu):
    """The as-written crossing (clamped secant, tol floor 0) on the 256^2 scene."""
    cfg, cam = K.preset("cpu256_a09")
    cfg.disk_root_mode = K.KT_DISK_ROOT_SECANT_REF
    cfg.tol_floor = 0.0
    cfg, cam = K.scaled(cfg, cam, 128, 128)
    say("[parity] fp64 secant-root tol_floor=0 128x128")
    g_hdr, g_cls, st = _render_gpu(torch_gpu, cfg, cam)
    o_hdr, o_cls, _, ost = O.render(cfg, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[parity] secant: {m}")
    assert m["class_mismatch_non_boundary"] == 0
    assert m["max_rel"] <= FP64_MAX_REL
    # bit for bit: every class, HDR float and attempt
    assert np.array_equal(g_cls, o_cls)
    assert np.array_equal(g_hdr[..., :3], o_hdr.astype(np.float32))
    assert int(st["attempts"]) == int(ost.attempts)


def test_reference_defaults_scene(torch_gpu):
    """config.yaml defaults (spin -0.45, M 0.5, spp 4, 100 steps) at 210x90."""
    cfg = K.default_config()
    cfg.width, cfg.height = 210, 90
    tree = {"pos": (80.0, 6.0, 0.0)}
    cam = K.camera_from_yaw_pitch(K.clamp_to_escape_radius(tree["pos"], cfg.escape_radius), -96.0, -4.0, 15.0)
    say("[parity] fp64 reference defaults 210x90 spp4")
    g_hdr, g_cls, st = _render_gpu(torch_gpu, cfg, cam)
    o_hdr, o_cls, _, ost = O.render(cfg, cam)
    m = compare(g_hdr, g_cls, o_hdr, o_cls)
    say(f"[parity] defaults: {m}")
    assert m["class_mismatch_non_boundary"] == 0
    assert m["class_mismatch_frac"] <= FP64_MAX_MISMATCH_FRAC
    assert m["max_rel"] <= FP64_MAX_REL
    # the as-written path (reference right-hand side, secant root, tolerance
    # floor 0, four samples per pixel) bit for bit as well
    assert np.array_equal(g_cls, o_cls)
    assert np.array_equal(g_hdr[..., :3], o_hdr.astype(np.float32))
    assert int(st["attempts"]) == int(ost.attempts)


def test_defaults_frame_matches_reference_artefact(torch_gpu):
    """The HIP frame of config.yaml's defaults (as-written modes, spp 4) +
    bloom/ACES/sRGB against the summary of the reference's first-frame
    artefact picture/diagram.png (check and thresholds: tests/diagram_check.py)."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    from tests import diagram_check as D
    cfg, cam = D.scene()
    say(f"[parity] defaults vs diagram.png summary {cfg.width}x{cfg.height}")
    with HipRenderer(cfg) as r:
        hdr, _ = r.trace(cam)
        ldr = r.post_process(hdr)
        torch.cuda.synchronize()
        g_ldr = ldr.cpu().numpy()
    m = D.metrics(g_ldr)
    say(f"[parity] diagram: {m}")
    assert D.passes(m), m


def test_fp64_frame_matches_golden_fixture(torch_gpu):
    """The 64x64 frame of the 256x256 CPU workload against the committed
    oracle frame (tests/golden/cpu256_a09_64x64.npz), without running the
    oracle: same classes and attempt counts per pixel, HDR to the FP64 bar."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    from tests.golden.make_oracle_golden import OUT, scene
    with np.load(OUT) as z:
        gold = {k: z[k] for k in z.files}
    cfg, cam = scene()
    say(f"[parity] golden {cfg.width}x{cfg.height}")
    with HipRenderer(cfg) as r:
        att = torch.zeros((cfg.height, cfg.width), dtype=torch.int32, device="cuda")
        hdr, cls = r.trace(cam, attempts=att)
        torch.cuda.synchronize()
        g_hdr, g_cls, g_att = hdr.cpu().numpy(), cls.cpu().numpy(), att.cpu().numpy()
    m = compare(g_hdr, g_cls, gold["hdr"], gold["cls"])
    say(f"[parity] golden: {m}")
    assert np.array_equal(g_cls, gold["cls"])
    assert np.array_equal(g_att.astype(np.uint32), gold["attempts"])
    assert m["max_rel"] <= FP64_MAX_REL


def test_post_process_matches_oracle(torch_gpu):
    """Bloom + ACES + sRGB of the GPU HDR frame, +-1 LSB of the FP64 oracle."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("cpu256_a09")
    say("[parity] post-process cpu256_a09")
    with HipRenderer(cfg) as r:
        hdr, _ = r.trace(cam, with_class=False)
        ldr = r.post_process(hdr)
        torch.cuda.synchronize()
        h = hdr.cpu().numpy()
        g_ldr = ldr.cpu().numpy()
    o_ldr, _ = O.post(cfg, h)
    d, hist = ldr_diff(g_ldr, o_ldr)
    say(f"[parity] post max byte diff {d} hist {hist}")
    assert d <= 1


def test_post_process_on_another_stream_and_bad_buffers(torch_gpu):
    """post_process(stream=s) with library-made bloom and image buffers
    gives the same image as on the current stream; wrongly shaped or typed
    caller buffers are refused before any launch."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("cpu256_a09")
    cfg, cam = K.scaled(cfg, cam, 160, 96)
    say("[parity] post-process on a side stream")
    with HipRenderer(cfg) as r:
        hdr, _ = r.trace(cam, with_class=False)
        ref = r.post_process(hdr)
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        for _ in range(3):
            got = r.post_process(hdr, stream=s)
            s.synchronize()
            assert torch.equal(got, ref)
        H, W = cfg.height, cfg.width
        bad = [dict(ldr=torch.empty((H, W - 1), dtype=torch.int32, device="cuda")),
               dict(ldr=torch.empty((H, W), dtype=torch.float32, device="cuda")),
               dict(bloom=torch.empty((H, W, 3), dtype=torch.float32, device="cuda")),
               dict(bloom=torch.empty((H, W, 4), dtype=torch.float64, device="cuda"))]
        for kw in bad:
            with pytest.raises(ValueError):
                r.post_process(hdr, **kw)
        with pytest.raises(ValueError):
            r.post_process(hdr[:, :-1])
        with pytest.raises(ValueError):
            r.autotune(cam, hdr=torch.empty((H, W, 4), dtype=torch.float16, device="cuda"))


def test_post_process_no_bloom_and_threshold(torch_gpu):
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    rng = np.random.default_rng(3)
    for enabled, thr, radius in [(0, 0.0, 24.0), (1, 0.0, 7.5), (1, 0.3, 36.0)]:
        cfg, cam = K.preset("cpu256_a09")
        cfg.width, cfg.height = 300, 77
        cfg.bloom_enabled, cfg.bloom_threshold, cfg.bloom_radius = enabled, thr, radius
        say(f"[parity] post synthetic enabled={enabled} thr={thr} R={radius}")
        h = np.ones((77, 300, 4), dtype=np.float32)
        h[..., :3] = (rng.random((77, 300, 3)) ** 3 * 4.0).astype(np.float32)
        with HipRenderer(cfg) as r:
            t = torch.from_numpy(h).cuda()
            g_ldr = r.post_process(t).cpu().numpy()
        o_ldr, _ = O.post(cfg, h)
        d, hist = ldr_diff(g_ldr, o_ldr)
        say(f"[parity]   max diff {d} hist {hist}")
        assert d <= 1


@pytest.mark.parametrize("w,h,radius", [(211, 37, 1.0), (300, 77, 2.5), (97, 230, 80.0), (120, 150, 90.0),
                                        (64, 260, 100.0), (70, 240, 104.0), (1, 1, 12.0), (2, 300, 36.0),
                                        (641, 3, 36.0)])
def test_post_process_tile_plans(torch_gpu, w, h, radius):
    """Bloom radii that take every vertical-pass tile plan (8 rows per lane
    with 8, 4, 2 and 1 waves per block, 1 row per lane for the widest
    radius), radii whose taps are fewer than a register tile, and frames
    narrower or shorter than a tile or than the blur: +-1 LSB of the FP64
    oracle."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    rng = np.random.default_rng(11)
    cfg, _ = K.preset("cpu256_a09")
    cfg.width, cfg.height = w, h
    cfg.bloom_radius = radius
    say(f"[parity] post {w}x{h} radius {radius}")
    hh = np.ones((h, w, 4), dtype=np.float32)
    hh[..., :3] = (rng.random((h, w, 3)) ** 3 * 4.0).astype(np.float32)
    with HipRenderer(cfg) as r:
        g_ldr = r.post_process(torch.from_numpy(hh).cuda()).cpu().numpy()
    o_ldr, _ = O.post(cfg, hh)
    d, hist = ldr_diff(g_ldr, o_ldr)
    say(f"[parity]   max diff {d} hist {hist}")
    assert d <= 1


def test_post_process_radius_beyond_lds_is_refused(torch_gpu):
    from kerrtrace_amd._lib import KerrTraceError
    from kerrtrace_amd.renderer import HipRenderer
    cfg, _ = K.preset("cpu256_a09")
    cfg.bloom_radius = 105.0
    with pytest.raises(KerrTraceError, match="bloom radius"):
        HipRenderer(cfg)


def test_frame_api_sequencing(torch_gpu):
    """submit/ready/finish semantics of src/renderer/mod.rs:119-260."""
    from kerrtrace_amd._lib import KerrTraceError
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("cpu256_a09")
    cfg, cam = K.scaled(cfg, cam, 64, 48)
    say("[parity] frame api 64x48")
    with HipRenderer(cfg) as r:
        assert r.submit_render_camera(cam)
        assert r.submit_render_camera(cam)
        assert not r.submit_render_camera(cam)  # both frames in flight
        torch_gpu.cuda.synchronize()
        import time
        f = None
        for _ in range(2000):
            f = r.ready_frame()
            if f is not None:
                break
            time.sleep(0.005)
        assert f is not None and f.index == 0
        first = f.data.copy()
        with pytest.raises(KerrTraceError) as ei:
            r.finish_frame(1)  # out of order
        assert ei.value.code == -5
        r.finish_frame(0)
        f2 = None
        for _ in range(2000):
            f2 = r.ready_frame()
            if f2 is not None:
                break
            time.sleep(0.005)
        assert f2 is not None and f2.index == 1
        assert np.array_equal(first, f2.data)  # same camera, deterministic
        r.finish_frame(1)
        with pytest.raises(KerrTraceError):
            r.finish_frame(0)  # empty queue
        # the frame through the f32 entry point equals the trace + post path
        pos = [float(v) for v in cam.pos]
        ok = r.submit_render(pos, list(cam.fwd), list(cam.rgt), list(cam.up), float(cam.fov_scale))
        assert ok


@pytest.mark.parametrize("n_streams", [2, 3, 4])
def test_traces_on_several_streams(torch_gpu, n_streams):
    """Consecutive traces take four sets of launch state in turn, so up to
    four traces issued on different streams run concurrently: every frame of
    an interleaved run equals the frame traced alone, class buffer included,
    and the counters read afterwards are those of the last trace."""
    torch = torch_gpu
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("hd1080_a0998")
    cfg, cam = K.scaled(cfg, cam, 240, 136)
    say(f"[parity] {n_streams} streams 240x136")
    with HipRenderer(cfg) as r:
        ref_hdr, ref_cls = r.trace(cam)
        torch.cuda.synchronize()
        ref_st = r.stats()
        streams = [torch.cuda.Stream() for _ in range(n_streams)]
        nb = n_streams + 1
        bufs = [(r.alloc_hdr(), torch.empty_like(ref_cls)) for _ in range(nb)]
        for k in range(2 * nb):  # more buffers than streams, more traces than launch-state sets
            hdr, cls = bufs[k % nb]
            s = streams[k % n_streams]
            if k >= nb:
                s.wait_stream(streams[(k - nb) % n_streams])  # buffer k % nb was last written there
            r.trace(cam, hdr=hdr, cls=cls, stream=s)
        torch.cuda.synchronize()
        st = r.stats()
        for hdr, cls in bufs:
            assert torch.equal(hdr, ref_hdr)
            assert torch.equal(cls, ref_cls)
        for key in ("rays", "attempts", "accepted", "rhs_evals", "crossings", "max_attempts_ray"):
            assert st[key] == ref_st[key], key


def test_lut_overflow_rejects_frame(torch_gpu):
    """A temperature above lut_max_temp fails the frame (kernel.cu:3-10, mod.rs:236-241)."""
    from kerrtrace_amd._lib import KerrTraceError
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("cpu256_a09")
    cfg, cam = K.scaled(cfg, cam, 64, 64)
    cfg.lut_max_temp = 2000.0
    say("[parity] lut overflow 64x64")
    with HipRenderer(cfg) as r:
        assert r.submit_render_camera(cam)
        torch_gpu.cuda.synchronize()
        import time
        err = None
        for _ in range(2000):
            try:
                f = r.ready_frame()
            except KerrTraceError as e:
                err = e
                break
            if f is not None:
                break
            time.sleep(0.005)
        assert err is not None and err.code == -4
        hdr, cls = r.trace(cam)
        st = r.stats()
        assert st["lut_overflow"] == 1 and st["overflow_temperature"] > 2000.0
        assert (cls.cpu().numpy() & 16).any()


def test_first_frame_png_from_frame_api(torch_gpu, tmp_path):
    """src/app.rs:219-237: the first ready frame of the frame loop saved as
    an RGBA PNG equals trace + post of the same camera, alpha 255."""
    from kerrtrace_amd import export as E
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("cpu256_a09")
    cfg, cam = K.scaled(cfg, cam, 96, 64)
    say("[parity] first-frame png 96x64")
    path = tmp_path / "picture" / "first_frame.png"
    with HipRenderer(cfg) as r:
        assert r.submit_render_camera(cam)
        import time
        f = None
        for _ in range(2000):
            f = r.ready_frame()
            if f is not None:
                break
            time.sleep(0.005)
        assert f is not None
        assert E.save_first_frame(str(path), f.data)
        r.finish_frame(f.index)
        hdr, _ = r.trace(cam, with_class=False)
        ldr = r.post_process(hdr)
        torch_gpu.cuda.synchronize()
        direct = ldr.cpu().numpy()
    assert not E.save_first_frame(str(path), direct)  # an existing file is kept
    from PIL import Image
    img = np.asarray(Image.open(str(path)))
    assert img.shape == (64, 96, 4) and np.all(img[..., 3] == 255)
    assert np.array_equal(img, E.ldr_to_rgba(direct))


@pytest.mark.parametrize("size", [(1, 1), (1, 37), (37, 1), (13, 7), (65, 3)])
def test_tiny_and_ragged_frames(torch_gpu, size):
    """Frames smaller than a tile and sizes that are no tile multiple, with
    tile edges 1, 8 and 16: every pixel traced once, bitwise equal to the
    oracle; the post-process with a bloom radius larger than the frame."""
    from kerrtrace_amd.renderer import HipRenderer
    cfg, cam = K.preset("hd1080_a0998")
    cfg, cam = K.scaled(cfg, cam, *size)
    say(f"[parity] ragged {size}")
    o_hdr, o_cls, o_att, ost = O.render(cfg, cam)
    with HipRenderer(cfg) as r:
        for chunk in (1, 8, 16):
            r.set_launch_opts(waves_per_cu=4, refill_threshold=8, queue_chunk=chunk, cost_order=1)
            att = torch_gpu.zeros(size[::-1], dtype=torch_gpu.int32, device="cuda")
            hdr, cls = r.trace(cam, attempts=att)
            ldr = r.post_process(hdr)
            torch_gpu.cuda.synchronize()
            g_hdr, g_cls = hdr.cpu().numpy(), cls.cpu().numpy()
            m = compare(g_hdr, g_cls, o_hdr, o_cls)
            assert m["class_mismatch"] == 0 and m["exact_fp32_frac"] == 1.0, (chunk, m)
            assert np.array_equal(att.cpu().numpy().astype(np.uint32), o_att), chunk
            st = r.stats()
            assert st["rays"] == size[0] * size[1] and st["attempts"] == ost.attempts
            o_ldr, _ = O.post(cfg, g_hdr)
            d, _ = ldr_diff(ldr.cpu().numpy(), o_ldr)
            assert d <= 1
