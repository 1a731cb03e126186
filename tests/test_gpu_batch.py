"""Batching layer on the GPU: HIP-graph replay, uneven multi-size batches and
the launch-shape autotuner, each checked against eager calls and the FP64
oracle (gate 1e-5 of clear-field intensity, aerial.rs:1682-1759)."""
import numpy as np
import pytest

import oracle
import highuvlith_amd as H

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _arf_engine(n, px, K, ctx=None):
    return H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                               H.ProjectionOptics.immersion_193i(), H.GridConfig(n, px),
                               max_kernels=K, context=ctx)


def _masks(n, count, seed):
    rng = np.random.default_rng(seed)
    m = np.zeros((count, n, n), np.float32)
    for c in range(count):
        for _ in range(10):
            w, h = rng.integers(n // 32, n // 4, 2)
            x, y = rng.integers(0, n - w), rng.integers(0, n - h)
            m[c, y:y + h, x:x + w] = 1.0
    return m


def _oracle(eng, mask, z=0.0):
    spec = oracle.fft2(mask.astype(np.complex128))
    return oracle.socs_image(eng.kernel_set(z), spec, eng.flare_fraction(), threads=8)


def test_graph_replays_match_eager_calls():
    import torch
    eng = _arf_engine(512, 2.0, 8)
    zs = (0.0, 40.0)
    planes = [eng.device_kernel_set(z) for z in zs]
    ctx = eng.context()
    masks = torch.from_numpy(_masks(512, 3, 1)).cuda()
    out_g = torch.empty((3, 2, 512, 512), device="cuda")
    out_e = torch.empty_like(out_g)
    for rep in range(4):
        # new contents in the same buffers, eager torch work between replays
        masks.copy_(torch.from_numpy(_masks(512, 3, 10 + rep)).cuda())
        scratch = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        H.engine.socs_images_device(ctx, masks, planes, out_g, graph=True)
        H.engine.socs_images_device(ctx, masks, planes, out_e, graph=False)
        torch.cuda.synchronize()
        diff = (out_g - out_e).abs().max().item()
        print("replay", rep, "max |graph - eager|", diff, float(scratch))
        assert diff == 0.0
    ref = _oracle(eng, masks[1].cpu().numpy(), 40.0)
    assert np.abs(out_g[1, 1].cpu().numpy() - ref).max() <= TOL


def test_clip_ranges_equal_one_range():
    """A call issued as several clip ranges (huv_socs_tuning.chunks: band,
    fields, coarse transforms and images per range, one range after the other)
    gives the same images bit for bit as one range: ragged last ranges, more
    ranges than clips, the fused (m <= 192) and two-pass (m = 256) coarse
    stage, real and complex rasters and rectangle lists, eager calls and graph
    replays with torch work on the same buffers between them."""
    import torch
    eng = _arf_engine(256, 2.0, 10)
    ctx = eng.context()
    planes = [eng.device_kernel_set(0.0), eng.device_kernel_set(40.0)]
    B, K = planes[0].band, max(p.host.num_kernels for p in planes)
    N = 7
    masks = torch.from_numpy(_masks(256, N, 21)).cuda()
    cmasks = (masks * torch.exp(1j * 0.3 * masks)).to(torch.complex64).contiguous()
    rng = np.random.default_rng(5)
    rect_clips = [[(float(rng.uniform(-200, 200)), float(rng.uniform(-200, 200)),
                    float(rng.uniform(20, 150)), float(rng.uniform(20, 150)), 1.0)
                   for _ in range(6)] for _ in range(N)]
    arrays = H.engine.rect_arrays(rect_clips)
    grid = H.GridConfig(256, 2.0)
    calls = {
        "real": lambda o, g: H.engine.socs_images_device(ctx, masks, planes, o, graph=g),
        "complex": lambda o, g: H.engine.socs_images_complex_device(ctx, cmasks, planes, o,
                                                                     graph=g),
        "rects": lambda o, g: H.engine.socs_images_rects_device(ctx, grid, arrays, planes, o),
    }
    ref = torch.empty((N, 2, 256, 256), device="cuda")
    out = torch.empty_like(ref)
    sizes = [m for m in (64, 128, 256) if m >= 4 * B + 1]
    assert 256 in sizes and len(sizes) >= 2
    for m in sizes:
        for kind, call in calls.items():
            if kind != "real" and m != sizes[0]:
                continue
            ctx.set_tuning(256, B, 2, K, m=m, chunks=1)
            call(ref, False)
            torch.cuda.synchronize()
            if kind == "real":
                err = np.abs(ref[6, 1].cpu().numpy()
                             - _oracle(eng, masks[6].cpu().numpy(), 40.0)).max()
                assert err <= TOL
            for chunks in (2, 3, 4, 16):
                ctx.set_tuning(256, B, 2, K, m=m, chunks=chunks)
                out.fill_(-1.0)
                call(out, False)
                torch.cuda.synchronize()
                assert torch.equal(out, ref), (kind, m, chunks)
                if kind == "rects":
                    continue
                for rep in range(4):  # the first call records, the others replay
                    out.fill_(-1.0)
                    call(out, True)
                    scratch = torch.randn(1 << 18, device="cuda").sum()
                    torch.cuda.synchronize()
                    print(kind, "m", m, "chunks", chunks, "replay", rep, float(scratch))
                    assert torch.equal(out, ref), (kind, m, chunks, rep)
    with pytest.raises(H.HuvError):
        ctx.set_tuning(256, B, 2, K, chunks=17)


def test_uneven_batch_matches_per_group_calls_and_oracle():
    import torch
    ctx = H.DeviceContext(0)
    groups, singles = [], []
    for n, count, K in [(256, 3, 12), (512, 2, 16), (1024, 1, 16)]:
        eng = _arf_engine(n, 2.0, K, ctx)
        planes = [eng.device_kernel_set(0.0)]
        masks = torch.from_numpy(_masks(n, count, n)).cuda()
        out = torch.empty((count, 1, n, n), device="cuda")
        groups.append((masks, planes, out))
        singles.append((eng, masks, planes))
    for graph in (False, True, True):
        for _, _, out in groups:
            out.zero_()
        H.engine.socs_images_batch_device(ctx, groups, graph=graph)
        torch.cuda.synchronize()
        for (eng, masks, planes), (_, _, out) in zip(singles, groups):
            ref_gpu = torch.empty_like(out)
            H.engine.socs_images_device(ctx, masks, planes, ref_gpu)
            torch.cuda.synchronize()
            assert (ref_gpu - out).abs().max().item() == 0.0
    for eng, masks, _ in singles:
        n = masks.shape[-1]
        out = [g for g in groups if g[0].shape[-1] == n][0][2]
        err = np.abs(out[0, 0].cpu().numpy() - _oracle(eng, masks[0].cpu().numpy())).max()
        print("uneven", n, "err", err)
        assert err <= TOL


def test_autotuner_records_a_valid_shape_within_the_gate():
    import torch
    eng = _arf_engine(1024, 2.0, 12)
    ctx = eng.context()
    ctx.set_autotune(True)
    planes = [eng.device_kernel_set(0.0)]
    B = planes[0].band
    assert ctx.tuning(1024, B, 1, planes[0].host.num_kernels) is None
    masks = torch.from_numpy(_masks(1024, 2, 5)).cuda()
    out = torch.empty((2, 1, 1024, 1024), device="cuda")
    H.engine.socs_images_device(ctx, masks, planes, out)
    torch.cuda.synchronize()
    t = ctx.tuning(1024, B, 1, planes[0].host.num_kernels)
    print("band", B, "tuned (m, n1, xnb, xlanes, chunks)", t)
    assert t is not None and t[0] >= 4 * B + 1 and t[0] in (32, 64, 128, 192, 256)
    assert t[1] in (16, 32) and t[2] in (4, 6, 8) and t[3] in (1, 2)
    for c in range(2):
        err = np.abs(out[c, 0].cpu().numpy() - _oracle(eng, masks[c].cpu().numpy())).max()
        assert err <= TOL
    # every forced shape stays within the gate too
    for m in (128, 256):
        for xlanes in (1, 2):
            ctx.set_tuning(1024, B, 1, planes[0].host.num_kernels, m=m, n1=16, xnb=4,
                           xlanes=xlanes)
            H.engine.socs_images_device(ctx, masks, planes, out)
            err = np.abs(out[1, 0].cpu().numpy() - _oracle(eng, masks[1].cpu().numpy())).max()
            print("forced m", m, "xlanes", xlanes, "err", err)
            assert err <= TOL
    with pytest.raises(H.HuvError):
        ctx.set_tuning(1024, B, 1, 12, m=64)   # below 4B + 1
    with pytest.raises(H.HuvError):
        ctx.set_tuning(1024, B, 1, 12, n1=8)   # no 8 x 128 factorisation
    with pytest.raises(H.HuvError):
        ctx.set_tuning(1024, B, 1, 12, xnb=5)
    with pytest.raises(H.HuvError):
        ctx.set_tuning(1024, B, 1, 12, xlanes=3)


def test_field_kernel_forms_agree_at_every_coarse_size():
    """Every x-pass form of the field kernel (one or two lanes per band row,
    each load block) against the one-lane reference form and the oracle, at
    each coarse grid the band allows; two lanes are refused below m = 64."""
    import torch
    eng = _arf_engine(256, 2.0, 10)
    ctx = eng.context()
    planes = [eng.device_kernel_set(0.0), eng.device_kernel_set(40.0)]
    B, K = planes[0].band, max(p.host.num_kernels for p in planes)
    masks = torch.from_numpy(_masks(256, 3, 11)).cuda()
    ref = torch.empty((3, 2, 256, 256), device="cuda")
    out = torch.empty_like(ref)
    sizes = [m for m in (32, 64, 128, 192, 256) if m >= 4 * B + 1]
    assert 32 in sizes or 64 in sizes
    for m in sizes:
        ctx.set_tuning(256, B, 2, K, m=m, xnb=6, xlanes=1)
        H.engine.socs_images_device(ctx, masks, planes, ref)
        for xlanes, xnb in ((1, 4), (1, 8), (2, 4), (2, 6), (2, 8)):
            if m == 32 and xlanes == 2:
                with pytest.raises(H.HuvError):
                    ctx.set_tuning(256, B, 2, K, m=m, xnb=xnb, xlanes=xlanes)
                continue
            ctx.set_tuning(256, B, 2, K, m=m, xnb=xnb, xlanes=xlanes)
            H.engine.socs_images_device(ctx, masks, planes, out)
            torch.cuda.synchronize()
            d = (out - ref).abs().max().item()
            err = np.abs(out[2, 0].cpu().numpy() - _oracle(eng, masks[2].cpu().numpy())).max()
            print("m", m, "xlanes", xlanes, "xnb", xnb, "vs one-lane", d, "vs oracle", err)
            assert d <= TOL and err <= TOL


def test_graph_is_dropped_with_its_kernel_sets():
    import torch
    ctx = H.DeviceContext(0)
    masks = torch.from_numpy(_masks(256, 2, 7)).cuda()
    out = torch.empty((2, 1, 256, 256), device="cuda")
    for K in (6, 10):
        eng = _arf_engine(256, 4.0, K, ctx)
        planes = [eng.device_kernel_set(0.0)]
        H.engine.socs_images_device(ctx, masks, planes, out, graph=True)
        H.engine.socs_images_device(ctx, masks, planes, out, graph=True)
        torch.cuda.synchronize()
        err = np.abs(out[0, 0].cpu().numpy() - _oracle(eng, masks[0].cpu().numpy())).max()
        print("K", K, "err", err)
        assert err <= TOL
        del planes, eng


def test_uneven_batch_at_k32_with_a_2048_class():
    """The uneven-batch workload's shape classes at its own kernel count: edges
    256/512/1024/2048 px at 2 nm, ArF 193i annular 0.7/0.9, K = 32 per size, one
    graph launch, 30-60 % dense Manhattan content.  The batch equals the
    per-group calls bit for bit (graph replays included) and the first clip of
    every class matches the FP64 oracle within the gate."""
    import torch
    from highuvlith_amd import workloads as W
    ctx = H.DeviceContext(0)
    rng = np.random.default_rng(17)
    groups, engs = [], []
    for n, count in [(256, 3), (512, 2), (1024, 2), (2048, 1)]:
        src, opt, grid, K = W.arf_annular_2nm_setup(n)
        eng = H.AerialImageEngine(src, opt, grid, max_kernels=K, context=ctx)
        assert eng.num_kernels() == 32
        planes = [eng.device_kernel_set(0.0)]
        cf = torch.from_numpy(rng.uniform(0.3, 0.6, count))
        masks = W.manhattan_clips(count, n, seed=200 + n, clear_fraction=cf,
                                  max_side=max(16, n // 8))
        out = torch.empty((count, 1, n, n), device="cuda")
        groups.append((masks, planes, out))
        engs.append(eng)
    for rep in range(3):
        for _, _, out in groups:
            out.fill_(float("nan"))
        H.engine.socs_images_batch_device(ctx, groups, graph=True)
        torch.cuda.synchronize()
        for masks, planes, out in groups:
            single = torch.empty_like(out)
            H.engine.socs_images_device(ctx, masks, planes, single)
            torch.cuda.synchronize()
            assert (single - out).abs().max().item() == 0.0, (rep, masks.shape)
    for eng, (masks, planes, out) in zip(engs, groups):
        n = masks.shape[-1]
        err = np.abs(out[0, 0].cpu().numpy() - _oracle(eng, masks[0].cpu().numpy())).max()
        print("uneven K=32", n, "band", planes[0].band, "coarse", planes[0].coarse, "err", err)
        assert err <= TOL
