"""The serving loop (vvr_server_*) on the GPU, through the C ABI.

- Each request's frames in a mix of ray-marched and path-traced requests on
  several volumes equal, bit for bit, the frames of the same request rendered
  alone by its renderer (path-traced ones with samples_per_frame = the
  server's increment, frame after frame): work items carry their request, each
  path its own seed, and each pixel's paths are summed in path order; also
  with a ghosted label in the mix and with path-traced requests in isolate mode
  only (alone, the renderer runs its ghost-free and density-free free flights).
- Progressive requests deliver frames in order, at most one per increment
  (increments that finish together deliver the newest), the last with every path.
- A small HBM budget makes requests wait for admission; all still finish and
  the budget is never exceeded once more than one request is admitted.
- Requests the renderer would reject are rejected at submission.
- Path-traced requests with the a-trous or the learned denoiser deliver
  denoised frames, equal to their renderer's.
- Cancelling ends a queued request at once, and a path-traced one in flight
  with its last delivered frame.
- A renderer destroyed while its requests are queued or running is deleted
  only when the last of them ends.
"""
import numpy as np
import pytest

from volview_amd import _abi, phantom, workloads
from volview_amd.renderer import Renderer, TransferFunction
from volview_amd.server import Server
from volview_amd.unet import Unet

pytestmark = pytest.mark.gpu


def volumes(ctx):
    hu_a = phantom.ct((48, 40, 36), 3)
    a = Renderer(ctx, hu_a, TransferFunction.preset(0))
    dims = (40, 48, 32)
    hu_b = phantom.ct(dims, 5)
    labels, soft, _, n = phantom.organs(dims, 5, 12)
    b = Renderer(ctx, hu_b, TransferFunction.preset(2))
    b.set_labels(labels, soft)
    return a, b, n


def mixed_requests(a, b, n, ghost=False):
    dvr_a = workloads.dvr_params(a, 0)
    pt_a = workloads.pt_params(a, 0, 8, bounces=4, camera_index=1)
    pt_learned = workloads.pt_params(a, 0, 8, bounces=4, camera_index=3)
    pt_learned.path_tracing.denoiser = _abi.DENOISE_LEARNED  # the network attached to `a`
    pt_b = workloads.pt_params(b, n, 12, bounces=4, camera_index=2)
    pt_b.path_tracing.denoiser = _abi.DENOISE_ATROUS  # each delivered frame denoised
    dvr_b = workloads.dvr_params(b, 1)
    workloads.organ_label_params(dvr_b, n, 7)  # same labels as pt_b: one opacity key per volume
    if ghost:  # while pt_b is in flight the free flights keep the ghost branch; alone, pt_a's do not
        for q in (pt_b, dvr_b):
            q.labels.ghost[2] = 1
            q.labels.ghost[5] = 1
    return [(a, dvr_a, 64, 48), (a, pt_a, 32, 24), (b, pt_b, 24, 32), (b, dvr_b, 40, 40),
            (a, workloads.dvr_params(a, 3), 17, 29), (a, pt_learned, 36, 20)]


def render_alone(r, p, w, h, increment):
    r.resize(w, h)
    if not p.path_tracing_enabled:
        r.record(1, p)
        return [r.read_hdr()]
    total = p.path_tracing.samples_per_frame
    q = _abi.FrameParams.from_buffer_copy(p)
    q.path_tracing.samples_per_frame = increment
    frames = []
    for f in range(total // increment):
        r.record(f + 1, q)
        frames.append(r.read_hdr())
    return frames


def drain(srv, ids, timeout_polls=600, cancelled=()):
    got = {i: [] for i in ids}
    finals = set()
    for _ in range(timeout_polls):
        for c in srv.poll(wait_ms=100):
            assert c.status == 0 or (c.id in cancelled and c.status == _abi.VVR_ECANCELED), c.status
            got[c.id].append(c)
            if c.final_frame:
                finals.add(c.id)
        if finals == set(ids):
            return got
    raise AssertionError(f"requests did not finish: {set(ids) - finals}")


@pytest.mark.parametrize("ghost", [False, True, "isolate only"])
def test_served_frames_equal_alone(ctx, ghost):
    a, b, n = volumes(ctx)
    net = Unet.pretrained(ctx)
    a.set_denoiser(net)
    reqs = mixed_requests(a, b, n, ghost is True)
    if ghost == "isolate only":  # every path-traced request in isolate mode: no density fetches
        reqs = [q for q in reqs if q[0] is b]
    inc = 4
    srv = Server(ctx, pt_increment=inc)
    ids = [srv.submit(r, p, w, h, tag=k) for k, (r, p, w, h) in enumerate(reqs)]
    got = drain(srv, ids)
    served = {i: srv.read_hdr(i, reqs[k][2], reqs[k][3]) for k, i in enumerate(ids)}
    for k, i in enumerate(ids):
        r, p, w, h = reqs[k]
        spp = [c.spp for c in got[i]]
        if p.path_tracing_enabled:  # increments that finish together deliver their newest frame
            assert spp == sorted(set(spp)) and all(x % inc == 0 for x in spp), spp
            assert spp[-1] == p.path_tracing.samples_per_frame, spp
        else:
            assert spp == [0]
        assert all(c.tag == k for c in got[i])
        srv.release(i)
    stats = srv.stats()
    srv.close()
    assert stats["pt_paths"] == sum(w * h * p.path_tracing.samples_per_frame for _, p, w, h in reqs
                                    if p.path_tracing_enabled)
    for k, i in enumerate(ids):
        r, p, w, h = reqs[k]
        alone = render_alone(r, p, w, h, inc)[-1]
        assert np.array_equal(served[i], alone), (k, np.abs(served[i] - alone).max())
        assert np.isfinite(alone).all() and np.abs(alone).max() > 0


def test_budget_queues_requests_and_all_finish(ctx):
    a, b, n = volumes(ctx)
    net = Unet.pretrained(ctx)
    a.set_denoiser(net)
    reqs = mixed_requests(a, b, n) * 3
    # The largest request: the denoised path-traced one, its images, two
    # increments' chunk sums in flight and the denoiser's four images.
    px = 32 * 24
    one = px * (8 + 48) + 2 * 4 * 3 * px * 16 + 4 * px * 16
    srv = Server(ctx, pt_increment=4, hbm_budget=2 * one)
    ids = [srv.submit(r, p, w, h) for r, p, w, h in reqs]
    drain(srv, ids)
    stats = srv.stats()
    assert stats["completed"] == len(ids) and stats["hbm_in_use"] == 0
    assert stats["hbm_peak"] <= 2 * one
    for i in ids:
        srv.release(i)
    srv.close()


def test_rejected_requests(ctx):
    a, _, _ = volumes(ctx)
    srv = Server(ctx, hbm_budget=1 << 20)
    p = workloads.pt_params(a, 0, 4)
    p.path_tracing.denoiser = _abi.DENOISE_LEARNED
    with pytest.raises(_abi.VvrError) as e:
        srv.submit(a, p, 16, 16)
    assert e.value.code == _abi.VVR_EINVAL
    with pytest.raises(_abi.VvrError) as e:  # needs more than the whole budget
        srv.submit(a, workloads.pt_params(a, 0, 4), 1024, 1024)
    assert e.value.code == _abi.VVR_ELIMIT
    with pytest.raises(_abi.VvrError):
        srv.read_hdr(12345, 4, 4)
    srv.close()


def test_cancel_queued_and_running(ctx):
    a, _, _ = volumes(ctx)
    srv = Server(ctx, pt_increment=4, max_requests=1)
    long_pt = workloads.pt_params(a, 0, 400, bounces=4)
    running = srv.submit(a, long_pt, 32, 32)
    queued = srv.submit(a, workloads.dvr_params(a, 0), 16, 16)
    first = None
    for _ in range(300):
        cs = srv.poll(wait_ms=100)
        first = next((c for c in cs if c.id == running), None)
        if first is not None:
            break
    assert first is not None and first.spp % 4 == 0 and 4 <= first.spp < 400 and not first.final_frame
    srv.cancel(queued)  # max_requests = 1 keeps it queued behind the running request
    srv.cancel(running)
    got = drain(srv, [running, queued], cancelled=(queued,))
    assert [c.status for c in got[queued]] == [_abi.VVR_ECANCELED]
    last = got[running][-1]
    delivered = [c.spp for c in got[running][:-1]]
    # The final completion repeats the last delivered frame (the increment in
    # flight is abandoned), unless the read-back of a frame was in flight.
    assert last.final_frame and last.status == 0 and 4 <= last.spp < 400 and last.spp % 4 == 0
    assert last.spp >= max(delivered, default=4)
    with pytest.raises(_abi.VvrError):
        srv.read_hdr(queued, 16, 16)
    img = srv.read_hdr(running, 32, 32)
    assert np.isfinite(img).all()
    srv.release(running)
    srv.release(queued)
    srv.close()


def test_renderer_destroyed_while_served(ctx):
    """A renderer destroyed while its requests are running or queued stays
    alive until the last of them ends: the running one delivers its frames,
    the queued one can still be cancelled, and nothing reads freed memory."""
    a, b, n = volumes(ctx)
    srv = Server(ctx, pt_increment=4, max_requests=1)
    p = workloads.pt_params(a, 0, 24, bounces=4)
    running = srv.submit(a, p, 32, 24)
    queued = srv.submit(b, workloads.dvr_params(b, 0), 16, 16)
    a.close()  # its request is in flight: deletion waits for it
    b.close()  # its request waits for admission
    srv.cancel(queued)
    got = drain(srv, [running, queued], cancelled=(queued,))
    last = got[running][-1]
    assert last.final_frame and last.status == 0 and last.spp == 24
    img = srv.read_hdr(running, 32, 24)
    assert np.isfinite(img).all() and np.abs(img[..., :3]).max() > 0
    srv.release(running)
    srv.release(queued)
    srv.close()


def test_renderer_calls_refused_while_held(ctx):
    """A request holds its renderer from submission to its final frame: record,
    resize and the set_* calls on it fail with VVR_ESTATE meanwhile (they would
    rewrite volumes, tables or images the server's kernels read), and work again
    once the request has ended."""
    a, _, _ = volumes(ctx)
    a.resize(16, 16)
    srv = Server(ctx, pt_increment=4, max_requests=1)
    rid = srv.submit(a, workloads.pt_params(a, 0, 64, bounces=4), 32, 24)
    labels = np.zeros((36, 40, 48), np.uint8)
    refused = {
        "record": lambda: a.record(1, workloads.dvr_params(a, 0)),
        "resize": lambda: a.resize(8, 8),
        "set_labels": lambda: a.set_labels(labels),
        "set_transfer_function": lambda: a.set_transfer_function(TransferFunction.preset(1)),
        "set_environment_map": lambda: a.set_environment_map(np.ones((4, 8, 3), np.float32)),
        "set_denoiser": lambda: a.set_denoiser(None),
    }
    for name, call in refused.items():
        with pytest.raises(_abi.VvrError) as e:
            call()
        assert e.value.code == _abi.VVR_ESTATE, name
    got = drain(srv, [rid])
    assert got[rid][-1].final_frame and got[rid][-1].spp == 64
    srv.release(rid)
    a.record(1, workloads.dvr_params(a, 0))  # the hold is gone
    a.set_transfer_function(TransferFunction.preset(1))
    srv.close()


def test_autotune_picks_a_shape_and_the_requests_still_render(ctx):
    """vvr_server_autotune over a small mix returns one of its candidate shapes;
    a server of that shape then renders the mix, each frame equal to the
    request rendered alone."""
    hu = phantom.ct((48, 48, 40), 3)
    r = Renderer(ctx, hu, TransferFunction.preset(0))
    reqs = []
    for k in range(3):
        p = workloads.pt_params(r, 0, 2 + k, bounces=4, camera_index=k)
        reqs.append((r, p, 32, 24))
    reqs.append((r, workloads.dvr_params(r, 1), 40, 32))
    shape = Server.autotune(ctx, reqs)
    assert shape["slots"] in (1 << 20, 2 << 20, 4 << 20, 8 << 20), shape
    assert shape["iterations"] in (1, 2, 4, 8), shape
    alone = []
    for k, (_, p, w, h) in enumerate(reqs):
        r.resize(w, h)
        r.record(k + 1, p)
        alone.append(r.read_hdr())
    srv = Server(ctx, **shape)
    ids = [srv.submit(rr, p, w, h) for rr, p, w, h in reqs]
    done = set()
    for _ in range(600):
        for c in srv.poll(wait_ms=100):
            assert c.status == 0
            if c.final_frame:
                done.add(c.id)
        if done == set(ids):
            break
    assert done == set(ids)
    for i, a in zip(ids, alone):
        assert np.array_equal(srv.read_hdr(i).view(np.uint32), a.view(np.uint32))
        srv.release(i)
    srv.close()
    r.close()
