"""Batched energy scan over many partial waves with uneven matrix sizes.

Workload: n_waves partial waves, each with nch channels drawn from
{2, 4, 8, 16, 32, 64} and nr mesh points from {40, 60, 80, 100} (seeded),
solved at n_energies energies uniform in [e_min, e_max]: every (wave, energy)
system C(E) X = B (N = nch nr, nch right-hand sides) is assembled on the
device, factored in complex FP32 with FP64 refinement (classic or FGMRES), and
reduced to R and S on the device (src/rmatrix_hp.F90:246-368 per system, the
reference's rmatrix call looped over energies).

Execution (``EnergyScan``):
* lanes: L HIP streams, each with its own LU handle and grow-only buffers
  sized for the largest system it runs; systems are assigned to lanes
  longest-processing-time first by their N^3 cost, so large and small
  systems of different waves run concurrently and the latency-bound panel
  of one system overlaps the GEMMs of others.  At most L cooperative panel
  kernels are in flight, which keeps them within the co-residence bound.
* graphs: each lane's systems are captured once, on the lane's stream, into
  chain graphs of at most NODE_BUDGET kernels and replayed on that stream,
  so one scan costs a few hundred graph launches instead of millions of
  kernel launches from the host (a system launches 10^2 - 10^3 kernels).
  Captured systems are factored without the look-ahead stream (a graph only
  spans its lane's stream); concurrency comes from the lanes.
* plan: per size class (N, nch) the factorization arithmetic (FP32 or
  split-BF16), the refinement kind and its launched step count, the
  trailing-update GEMM tile and the look-ahead switch come from ``autotune``
  (measured on a few energies of that class), as does the number of lanes.
Nothing synchronises with the host inside a scan; S of every system is left
in device memory (``s_of(wave, energy)``).
"""
import ctypes
import json
import os
import time

import numpy as np
import torch

from . import _lib, rmatrix, solver, synth

NC_CHOICES = (2, 4, 8, 16, 32, 64)
NR_CHOICES = (40, 60, 80, 100)


def draw_waves(n_waves=64, seed=0):
    """(nch, nr) of each wave, drawn uniformly from the choice sets."""
    rng = np.random.default_rng(seed)
    nc = rng.choice(NC_CHOICES, n_waves)
    nr = rng.choice(NR_CHOICES, n_waves)
    return [(int(a), int(b)) for a, b in zip(nc, nr)]


def _dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


class Wave:
    """One partial wave: synthetic system (pair-recipe potential, thresholds
    in [0, emax]) and its device-resident per-energy channel data."""

    def __init__(self, index, nch, nr, energies, a=15.0, emax=3.0, seed=0):
        self.index, self.nch, self.nr, self.n, self.a = index, nch, nr, nch * nr, a
        self.sysm = synth.System(nch, nr, a, emax, float(energies[0]), seed + 7919 * (index + 1))
        self.dmesh = rmatrix.DeviceMesh(self.sysm.mesh)
        self.lv = _dev(np.asarray(self.sysm.lval, dtype=np.int32), torch.int32)
        self.cpot = _dev(self.sysm.cpot, torch.complex128)
        self.b = rmatrix.rhs_colmajor(self.dmesh, nch)
        self.energies = np.asarray(energies, dtype=np.float64)
        ne = len(self.energies)
        qk = np.stack([synth.wave_numbers(self.sysm.eps, e) for e in self.energies])
        self.qk_host = qk
        self.qk = _dev(qk, torch.float64)                       # (ne, nch)
        xc1 = np.zeros((ne, nch))
        opn = np.zeros((ne, nch), dtype=np.int32)
        co = np.zeros((ne, nch), dtype=np.complex128)
        cop = np.zeros((ne, nch), dtype=np.complex128)
        self.nopen = np.zeros(ne, dtype=np.int64)
        for k in range(ne):
            o, c0, c1, x1 = rmatrix.channel_data(self.sysm.lval, qk[k], a)
            xc1[k] = x1
            self.nopen[k] = len(o)
            opn[k, :len(o)] = o
            co[k, :len(o)] = c0
            cop[k, :len(o)] = c1
        self.xc1 = _dev(xc1, torch.float64)
        self.open = _dev(opn, torch.int32)
        self.co = _dev(co, torch.complex128)
        self.cop = _dev(cop, torch.complex128)
        self.nopen_dev = _dev(self.nopen.astype(np.int32), torch.int32)
        self.R = torch.zeros((ne, nch, nch), dtype=torch.complex128, device='cuda')
        self.S = torch.zeros((ne, nch, nch), dtype=torch.complex128, device='cuda')
        # per-energy outcome of the batched solve: relative residual and status
        # (1 converged, 0 not converged, -info exactly singular)
        self.res = torch.zeros(ne, dtype=torch.float64, device='cuda')
        self.status = torch.zeros(ne, dtype=torch.int32, device='cuda')

    def s_of(self, k):
        """S at energy index k (nopen x nopen matrix, from the column-major slot)."""
        no = int(self.nopen[k])
        return self.S[k].reshape(-1)[:no * no].reshape(no, no).transpose(0, 1)

    @property
    def cost(self):
        return float(self.n) ** 3


class Lane:
    """One stream with its own LU handle, matrix/RHS buffers and R->S workspace."""

    def __init__(self, max_n, max_nch):
        lib = _lib.load()
        self.stream = torch.cuda.Stream()
        self.lu = solver.LU()
        self.a = torch.empty(max_n * max_n, dtype=torch.complex128, device='cuda')
        self.x = torch.empty(max_nch * max_n, dtype=torch.complex128, device='cuda')
        self.work_bytes = int(lib.hprmat_hip_r_to_s_work_bytes(max_nch))
        self.work = torch.empty(self.work_bytes, dtype=torch.uint8, device='cuda')
        self.items = []


DEFAULT_PLAN = {'method': 'fp32', 'refinement': 'fgmres', 'max_steps': 16, 'trail': 0,
                'lookahead': 1}


def solve_system(lane, wave, k, plan, tol=1e-12, stats=False):
    """Assemble, factor, refine and reduce system (wave, k) on the current
    stream; with stats=True returns the refinement statistics (synchronises)."""
    lib = _lib.load()
    n, nch, nr = wave.n, wave.nch, wave.nr
    st = solver._stream_ptr()
    a = lane.a[:n * n].view(n, n)
    dm = wave.dmesh
    rc = lib.hprmat_hip_assemble(
        nch, nr, ctypes.c_void_p(dm.tc.data_ptr()), ctypes.c_void_p(dm.blo0.data_ptr()),
        ctypes.c_void_p(dm.r.data_ptr()), ctypes.c_double(wave.a), ctypes.c_void_p(wave.lv.data_ptr()),
        ctypes.c_void_p(wave.qk[k].data_ptr()), ctypes.c_void_p(wave.xc1[k].data_ptr()),
        ctypes.c_void_p(wave.cpot.data_ptr()), ctypes.c_void_p(a.data_ptr()), ctypes.c_int64(n), st)
    _lib.check(rc, 'hprmat_hip_assemble')
    x = lane.x[:nch * n].view(nch, n)
    x.copy_(wave.b)
    lane.lu.set_plan(trail_tile=plan['trail'], lookahead=plan['lookahead'])
    lane.lu.factor(a, plan['method'], equilibrate=True)
    st_ref = lane.lu.solve_(x, plan['max_steps'], tol, 'rel_fro', True, stats=stats,
                            refinement=plan['refinement'])
    no = int(wave.nopen[k])
    rc = lib.hprmat_hip_r_to_s_ws(
        nch, nr, ctypes.c_void_p(x.data_ptr()), ctypes.c_int64(n), ctypes.c_void_p(dm.q2.data_ptr()),
        ctypes.c_double(1.0), ctypes.c_double(wave.a), ctypes.c_void_p(wave.qk[k].data_ptr()), no,
        ctypes.c_void_p(wave.open[k].data_ptr()), ctypes.c_void_p(wave.co[k].data_ptr()),
        ctypes.c_void_p(wave.cop[k].data_ptr()), ctypes.c_void_p(wave.R[k].data_ptr()),
        ctypes.c_void_p(wave.S[k].data_ptr()), ctypes.c_void_p(lane.work.data_ptr()),
        ctypes.c_size_t(lane.work_bytes), st)
    _lib.check(rc, 'hprmat_hip_r_to_s_ws')
    return st_ref


def _lpt(items, nlanes):
    """Longest-processing-time-first assignment of (cost, payload) items."""
    loads = [0.0] * nlanes
    out = [[] for _ in range(nlanes)]
    for cost, it in sorted(items, key=lambda t: -t[0]):
        j = int(np.argmin(loads))
        loads[j] += cost
        out[j].append(it)
    return out


NODE_BUDGET = 60000     # kernel nodes per captured graph


def est_nodes(w, plan):
    """Upper estimate of the kernels one system launches (the refinement
    launches all max_steps steps; the device skips those after convergence)."""
    nb = (w.n + 255) // 256
    return 30 + 50 * nb + plan['max_steps'] * (25 + 3 * nb)


class EnergyScan:
    def __init__(self, waves, lanes=4, graph_max=10 ** 9, plans=None, log=None):
        self.waves = waves
        self.nlanes = lanes
        self.graph_max = graph_max
        self.plans = plans or {}
        self.log = log
        items = [(w.cost, (w, k)) for w in waves for k in range(len(w.energies))]
        self.lanes = self._make_lanes(_lpt(items, lanes)) if items else []
        self.captured = False

    @staticmethod
    def _make_lanes(assign):
        lanes = []
        for items in assign:
            if not items:
                continue
            lane = Lane(max(w.n for w, _ in items), max(w.nch for w, _ in items))
            lane.items = sorted(items, key=lambda t: (-t[0].n, -t[0].nch))
            lane.graphs = []
            lanes.append(lane)
        return lanes

    def plan_for(self, w):
        p = self.plans.get((w.n, w.nch), DEFAULT_PLAN)
        if w.n <= self.graph_max:
            # graph-captured systems stay on their lane's stream: no look-ahead
            # (its second stream and reused events are kept out of captures;
            # the arithmetic is identical, the split falls on tile boundaries)
            p = dict(p, lookahead=0)
        return p

    def _say(self, msg):
        if self.log:
            print(msg, file=self.log, flush=True)

    def warm(self):
        """Eager pass that brings every handle's grow-only buffers (factors,
        right-hand sides, refinement workspace) to their final size and runs
        every kernel variant of the lane's plans once, so that nothing is
        allocated or set up for the first time inside a graph capture."""
        for lane in self.lanes:
            reps = {}
            first = lane.items[0]
            widest = max(lane.items, key=lambda t: t[0].n * t[0].nch)
            reps[id(first[0])] = first
            reps[id(widest[0])] = widest
            seen = set()
            for w, k in lane.items:
                key = tuple(sorted(self.plan_for(w).items()))
                if key not in seen:
                    seen.add(key)
                    reps.setdefault(id(w), (w, k))
            with torch.cuda.stream(lane.stream):
                for w, k in reps.values():
                    solve_system(lane, w, k, self.plan_for(w))
        torch.cuda.synchronize()

    def capture(self):
        """Capture each lane's systems (those up to N = graph_max) into chain
        graphs of at most NODE_BUDGET kernels, captured on the lane's stream."""
        self._say('# scan: warm-up')
        self.warm()
        total = 0
        for li, lane in enumerate(self.lanes):
            chunks, cur, nodes = [], [], 0
            for w, k in lane.items:
                if w.n > self.graph_max:
                    continue
                c = est_nodes(w, self.plan_for(w))
                if cur and nodes + c > NODE_BUDGET:
                    chunks.append(cur)
                    cur, nodes = [], 0
                cur.append((w, k))
                nodes += c
            if cur:
                chunks.append(cur)
            lane.graphs = []
            for chunk in chunks:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=lane.stream, capture_error_mode='relaxed'):
                    for w, k in chunk:
                        solve_system(lane, w, k, self.plan_for(w))
                lane.graphs.append(g)
            total += len(lane.graphs)
            self._say('# scan: lane %d: %d systems in %d graphs' % (li, len(lane.items),
                                                                    len(lane.graphs)))
        torch.cuda.synchronize()
        self.captured = True
        self._say('# scan: %d graphs captured' % total)

    def run(self):
        """One full scan: on every lane, its graphs replayed (or, before
        capture, its systems launched); lanes run concurrently.  Returns after
        enqueueing."""
        root = torch.cuda.current_stream()
        for lane in self.lanes:
            lane.stream.wait_stream(root)
            with torch.cuda.stream(lane.stream):
                if self.captured:
                    for w, k in lane.items:
                        if w.n > self.graph_max:
                            solve_system(lane, w, k, self.plan_for(w))
                    for g in lane.graphs:
                        g.replay()
                else:
                    for w, k in lane.items:
                        solve_system(lane, w, k, self.plan_for(w))
        for lane in self.lanes:
            root.wait_stream(lane.stream)

    @property
    def n_systems(self):
        return sum(len(w.energies) for w in self.waves)

    @property
    def flops(self):
        """Algorithmic flops of the scan: sum of 8/3 N^3 + 8 N^2 nrhs."""
        return float(sum(len(w.energies) * (8.0 / 3.0 * w.n ** 3 + 8.0 * w.n ** 2 * w.nch)
                         for w in self.waves))


class BatchScan:
    """Energy scan as uniform batches: the energies of one wave are systems
    of one order N, so they are assembled, factored, solved, refined and
    reduced to R and S in lockstep (hprmat_hip_assemble_batch,
    hprmat_hip_batch_solve, hprmat_hip_r_to_s_batch): one chain of launches
    per chunk of energies instead of one chain per system.  A chunk is as
    many energies as fit the lane's share of ``mem_bytes``.  With
    ``mesh_operator`` the refinement's FP64 products are formed from the
    Lagrange-mesh parts of the matrices (hprmat_hip_batch_solve_mesh).

    Plans: ``plans`` maps a size class (N, nch) to its factorization
    arithmetic ('fp32' / 'fp32_3m'), tile plan ('trail', 'subst': the
    HPRMAT_BATCH_TRAIL / _SUBST tiles) and launched sweep count
    ('max_refine'), as ``autotune_batch`` measures them; classes without an
    entry use ``method`` / ``max_refine`` and the default tiles.

    Waves are spread over ``lanes`` streams longest-processing-time first, so
    the launch-latency-bound chains of small waves overlap the MFMA-bound
    chains of large ones.  ``capture()`` records the whole scan as one HIP
    graph with the lanes as parallel branches (one graph launch per scan).  Refinement: FP64 classic sweeps
    (converged systems skip the rest).  ``finish()`` synchronises and
    re-solves any system left unconverged with the plain FP64 batch (the
    reference's fallback to its FP64 path when the fast one fails,
    src/rmat_solvers.F90:576-588).
    """

    def __init__(self, waves, method='fp32', lanes=2, mem_bytes=48 << 30, max_refine=20,
                 tol=1e-12, log=None, mesh_operator=True, plans=None):
        self.waves, self.method, self.max_refine, self.tol = waves, method, max_refine, tol
        self.mesh_operator = mesh_operator
        self.log = log
        self.captured = False
        self.graph = None          # the whole scan as one HIP graph, once captured
        self.fallbacks = 0
        self.plans = {}
        for w in waves:
            p = dict((plans or {}).get((w.n, w.nch), {}))
            p.setdefault('method', method)
            p.setdefault('trail', 0)
            p.setdefault('subst', 0)
            p.setdefault('max_refine', max_refine)
            self.plans[id(w)] = p
        budget = mem_bytes // max(1, lanes)
        self.chunk = {}
        for w in waves:
            per = (16 * w.n * w.n + 16 * w.n * w.nch
                   + solver.batch_work_bytes(self.plans[id(w)]['method'], w.n, w.nch, 1))
            self.chunk[id(w)] = int(max(1, min(len(w.energies), budget // per)))
        items = [(self._cost(w), w) for w in waves if len(w.energies)]
        self.lanes = []
        for ws in _lpt(items, lanes):
            if not ws:
                continue
            lane = _BatchLane()
            lane.waves = sorted(ws, key=lambda w: -w.n)
            lane.calls = [(w, k0, min(self.chunk[id(w)], len(w.energies) - k0))
                          for w in lane.waves for k0 in range(0, len(w.energies), self.chunk[id(w)])]
            ma = max(c * w.n * w.n for w, _, c in lane.calls)
            mx = max(c * w.n * w.nch for w, _, c in lane.calls)
            mw = max(solver.batch_work_bytes(self.plans[id(w)]['method'], w.n, w.nch, c)
                     for w, _, c in lane.calls)
            lane.a = torch.empty(ma, dtype=torch.complex128, device='cuda')
            lane.x = torch.empty(mx, dtype=torch.complex128, device='cuda')
            lane.work = torch.empty(mw, dtype=torch.uint8, device='cuda')
            self.lanes.append(lane)

    def plan_of(self, w):
        """The plan the scan runs wave w with."""
        return dict(self.plans[id(w)])

    def _cost(self, w):
        """Estimated seconds of a wave: MFMA work plus the launch chain."""
        ne, c = len(w.energies), self.chunk[id(w)]
        chunks = (ne + c - 1) // c
        return ne * (8.0 / 3.0) * w.n ** 3 / 60e12 + chunks * self._nodes(w, c) * 4e-6

    def _nodes(self, w, count):
        """Upper estimate of the kernels one chunk launches."""
        nb = (w.n + 255) // 256
        g = (w.n + 255) // 256
        leaf = 8 * ((count * g + 479) // 480)
        return 20 + nb * (45 + leaf) + (self.plans[id(w)]['max_refine'] + 1) * (10 + 3 * nb)

    def _say(self, msg):
        if self.log:
            print(msg, file=self.log, flush=True)

    def _chunk_call(self, lane, w, k0, c):
        n, nch = w.n, w.nch
        a = lane.a[:c * n * n].view(c, n, n)
        x = lane.x[:c * nch * n].view(c, nch, n)
        e = slice(k0, k0 + c)
        rmatrix.assemble_batch(w.dmesh, w.lv, w.qk[e], w.xc1[e], w.cpot, a)
        x.copy_(w.b.unsqueeze(0).expand(c, -1, -1))
        mesh = (solver.mesh_parts(w.dmesh, w.lv, w.qk[e], w.xc1[e], w.cpot)
                if self.mesh_operator else None)
        p = self.plans[id(w)]
        code = solver.batch_code(p['method'], trail=p['trail'], subst=p['subst'])
        solver.batch_solve(code, n, nch, c, a.data_ptr(), x.data_ptr(), p['max_refine'],
                           self.tol, 'rel_fro', w.res[k0].data_ptr(), w.status[k0].data_ptr(),
                           lane.work, mesh=mesh)
        rmatrix.r_to_s_batch(w.dmesh, x, w.qk[e], w.nopen_dev[e], w.open[e], w.co[e], w.cop[e],
                             w.R[e], w.S[e])

    def _lane_calls(self, lane, calls):
        with torch.cuda.stream(lane.stream):
            for w, k0, c in calls:
                self._chunk_call(lane, w, k0, c)

    def capture(self, single=True):
        """Eager warm-up (first-use kernel attributes), then the graphs.
        ``single`` (and at most NODE_BUDGET kernels in the whole scan): the
        scan as one HIP graph -- a capture stream forks to every lane's
        stream, each lane records its chunks, and the lanes join back, so the
        lanes become parallel branches of one graph and one graph launch
        replays the scan.  Otherwise (or if that capture fails) each lane's
        chunks go into chain graphs of at most NODE_BUDGET kernels replayed on
        the lane's stream.  (A capture of the 64-wave x 256-energy scan as one graph,
        about 2.4e5 kernels in three dependency chains of ~8e4, ended the
        process with a segmentation fault on the host, while chains of up to
        NODE_BUDGET kernels are captured and replayed routinely.)"""
        self._say('# batch scan: warm-up')
        self.run()
        torch.cuda.synchronize()
        self.graph = None
        nodes = sum(self._nodes(w, c) for lane in self.lanes for w, _, c in lane.calls)
        if single and nodes <= NODE_BUDGET:
            try:
                self._capture_single()
                self.captured = True
                return
            except RuntimeError as exc:
                self._say('# batch scan: single-graph capture failed (%s); lane graphs' % exc)
                self.graph = None
                torch.cuda.synchronize()
        total = 0
        for li, lane in enumerate(self.lanes):
            groups, cur, nodes = [], [], 0
            for call in lane.calls:
                c = self._nodes(call[0], call[2])
                if cur and nodes + c > NODE_BUDGET:
                    groups.append(cur)
                    cur, nodes = [], 0
                cur.append(call)
                nodes += c
            if cur:
                groups.append(cur)
            lane.graphs = []
            for grp in groups:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=lane.stream, capture_error_mode='relaxed'):
                    for w, k0, c in grp:
                        self._chunk_call(lane, w, k0, c)
                lane.graphs.append(g)
            total += len(lane.graphs)
            self._say('# batch scan: lane %d: %d waves, %d chunks, %d graphs'
                      % (li, len(lane.waves), len(lane.calls), len(lane.graphs)))
        torch.cuda.synchronize()
        self.captured = True
        self._say('# batch scan: %d graphs captured' % total)

    def _capture_single(self):
        root = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        t0 = time.perf_counter()
        with torch.cuda.graph(g, stream=root, capture_error_mode='relaxed'):
            for lane in self.lanes:
                lane.stream.wait_stream(root)            # fork
                self._lane_calls(lane, lane.calls)
                root.wait_stream(lane.stream)            # join
        torch.cuda.synchronize()
        self.graph, self.graph_stream = g, root
        self._say('# batch scan: one graph of %d lanes, %d chunks, captured in %.1f s'
                  % (len(self.lanes), sum(len(l.calls) for l in self.lanes),
                     time.perf_counter() - t0))

    def run(self):
        """Enqueue one full scan (one graph launch, or the lanes' graph replays
        once captured; eager launches before); returns without synchronising."""
        root = torch.cuda.current_stream()
        if self.captured and getattr(self, 'graph', None) is not None:
            self.graph_stream.wait_stream(root)
            with torch.cuda.stream(self.graph_stream):
                self.graph.replay()
            root.wait_stream(self.graph_stream)
            return
        for lane in self.lanes:
            lane.stream.wait_stream(root)
            if self.captured:
                with torch.cuda.stream(lane.stream):
                    for g in lane.graphs:
                        g.replay()
            else:
                self._lane_calls(lane, lane.calls)
        for lane in self.lanes:
            root.wait_stream(lane.stream)

    def finish(self):
        """Synchronise; re-solve every system whose status is not 1 with the
        plain FP64 batch (count 1, up to 10 sweeps).  Returns the number of
        systems re-solved; statuses still not 1 afterwards are left in
        ``wave.status``."""
        torch.cuda.synchronize()
        redo = []
        for w in self.waves:
            bad = torch.nonzero(w.status != 1).flatten().tolist()
            redo += [(w, k) for k in bad]
        self.fallbacks = len(redo)
        if not redo:
            return 0
        code = solver.batch_code('fp64')
        for w, k in redo:
            n, nch = w.n, w.nch
            a = torch.empty((1, n, n), dtype=torch.complex128, device='cuda')
            x = w.b.unsqueeze(0).clone()
            work = torch.empty(solver.batch_work_bytes('fp64', n, nch, 1), dtype=torch.uint8,
                               device='cuda')
            e = slice(k, k + 1)
            rmatrix.assemble_batch(w.dmesh, w.lv, w.qk[e], w.xc1[e], w.cpot, a)
            solver.batch_solve(code, n, nch, 1, a.data_ptr(), x.data_ptr(), 10, self.tol,
                               'rel_fro', w.res[k].data_ptr(), w.status[k].data_ptr(), work)
            rmatrix.r_to_s_batch(w.dmesh, x, w.qk[e], w.nopen_dev[e], w.open[e], w.co[e],
                                 w.cop[e], w.R[e], w.S[e])
        torch.cuda.synchronize()
        return len(redo)

    @property
    def n_systems(self):
        return sum(len(w.energies) for w in self.waves)

    @property
    def flops(self):
        """Algorithmic flops of the scan: sum of 8/3 N^3 + 8 N^2 nrhs."""
        return float(sum(len(w.energies) * (8.0 / 3.0 * w.n ** 3 + 8.0 * w.n ** 2 * w.nch)
                         for w in self.waves))


class _BatchLane:
    """One stream of a BatchScan with its matrix, solution and workspace
    buffers, sized for the largest chunk it runs."""

    def __init__(self):
        self.stream = torch.cuda.Stream()
        self.graphs = []


def serial_fp64(waves, pairs, tol=1e-12):
    """Reference loop: one plain FP64 solve per system, one stream, in order
    (factor + solve + FP64 refinement to the residual gate, at most two
    sweeps, as the plain method does; then R and S)."""
    lane = Lane(max(w.n for w, _ in pairs), max(w.nch for w, _ in pairs))
    plan = {'method': 'fp64', 'refinement': 'ir', 'max_steps': 2, 'trail': 0, 'lookahead': 1}
    for w, k in pairs:
        solve_system(lane, w, k, plan, tol)


def autotune(waves, sample=3, cache=None, lane_options=(2, 4, 8), log=None):
    """Pick, per size class (N, nch), the refinement kind, the trailing GEMM
    tile and the look-ahead switch by timing `sample` energies of the class on
    one lane; then the number of lanes by timing a slice of the whole scan.
    Results are cached in `cache` (JSON) when given."""
    if cache and os.path.exists(cache):
        with open(cache) as f:
            d = json.load(f)
        plans = {tuple(map(int, key.split('x'))): v for key, v in d['plans'].items()}
        return plans, int(d['lanes'])
    classes = {}
    for w in waves:
        classes.setdefault((w.n, w.nch), w)
    candidates = []
    for method, refinement in (('fp32', 'ir'), ('fp32', 'fgmres'), ('bf16x3', 'fgmres')):
        for trail in (0, 1, 2):
            for lookahead in (1, 0):
                if method == 'bf16x3' and trail != 0:
                    continue      # the tile applies to the 4-product kernels only
                candidates.append({'method': method, 'refinement': refinement,
                                   'max_steps': 16, 'trail': trail, 'lookahead': lookahead})
    plans = {}
    for key, w in sorted(classes.items()):
        if log:
            print('# autotune: class N=%d nch=%d' % key, file=log, flush=True)
        lane = Lane(w.n, w.nch)
        ks = list(range(0, len(w.energies), max(1, len(w.energies) // sample)))[:sample]
        best = None
        for p in candidates:
            if p['lookahead'] == 0 and w.n <= 512:
                continue          # no look-ahead below two panels anyway
            if p['trail'] != 0 and w.n <= 512:
                continue          # a single trailing update: the default tile
            with torch.cuda.stream(lane.stream):
                sts = [solve_system(lane, w, k, p, stats=True) for k in ks]
                if not all(st['converged'] for st in sts):
                    continue      # must reach the tolerance on every sampled energy
                # launched steps: the most the sample needed plus a margin
                p = dict(p, max_steps=min(16, max(st['sweeps'] for st in sts) + 3))
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for k in ks:
                    solve_system(lane, w, k, p)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, p)
        if best is None:
            best = (0.0, dict(DEFAULT_PLAN, max_steps=24))
        plans[key] = dict(best[1])
        del lane
    # lanes: time a slice (every 16th energy of every wave)
    best_l = None
    for nl in lane_options:
        if log:
            print('# autotune: %d lanes' % nl, file=log, flush=True)
        sc = EnergyScan(_slice_waves(waves, 16), lanes=nl, graph_max=0, plans=plans)
        sc.warm()
        t0 = time.perf_counter()
        sc.run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if best_l is None or dt < best_l[0]:
            best_l = (dt, nl)
        del sc
    if cache:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, 'w') as f:
            json.dump({'plans': {'%dx%d' % k: v for k, v in plans.items()}, 'lanes': best_l[1]}, f)
    return plans, best_l[1]


BATCH_REFINE_LADDER = (3, 4, 6, 8, 10, 12, 15, 20)


def _batch_trial(w, ks, plan, bufs, tol, mesh_operator, timed):
    """Solve the energies ``ks`` of wave ``w`` as one uniform batch with
    ``plan`` (eager, on the current stream).  Returns (every system
    converged, milliseconds of a second identical run when ``timed``)."""
    n, nch, c = w.n, w.nch, len(ks)
    a = bufs['a'][:c * n * n].view(c, n, n)
    x = bufs['x'][:c * nch * n].view(c, nch, n)
    idx = torch.as_tensor(np.asarray(ks), device='cuda')
    qk, xc1 = w.qk[idx].contiguous(), w.xc1[idx].contiguous()
    res = torch.zeros(c, dtype=torch.float64, device='cuda')
    status = torch.zeros(c, dtype=torch.int32, device='cuda')
    code = solver.batch_code(plan['method'], trail=plan['trail'], subst=plan['subst'])

    def once():
        rmatrix.assemble_batch(w.dmesh, w.lv, qk, xc1, w.cpot, a)
        x.copy_(w.b.unsqueeze(0).expand(c, -1, -1))
        mesh = solver.mesh_parts(w.dmesh, w.lv, qk, xc1, w.cpot) if mesh_operator else None
        solver.batch_solve(code, n, nch, c, a.data_ptr(), x.data_ptr(), plan['max_refine'], tol,
                           'rel_fro', res.data_ptr(), status.data_ptr(), bufs['work'], mesh=mesh)

    once()
    torch.cuda.synchronize()
    ok = bool((status == 1).all())
    if not timed:
        return ok, None
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    once()
    e1.record()
    torch.cuda.synchronize()
    return ok, e0.elapsed_time(e1)


def autotune_batch(waves, sample=16, mem_bytes=48 << 30, lane_options=(2, 3, 4), tol=1e-12,
                   mesh_operator=True, margin=4, min_refine=15, cache=None, log=None):
    """Plans of a BatchScan, measured on this GPU.  Per size class (N, nch),
    on ``sample`` energies spread over one wave of the class, solved as one
    batch: for each FP32 arithmetic ('fp32', 'fp32_3m') the fewest launched
    sweeps (BATCH_REFINE_LADDER) at which every sampled system meets ``tol``
    (bisection), timed; the faster arithmetic is kept and every tile plan
    (HPRMAT_BATCH_TRAIL x HPRMAT_BATCH_SUBST) is timed with it; the plan keeps
    the fastest tiles; its launched sweep count is the sampled one plus
    ``margin``, and at least ``min_refine``: a converged system skips the
    remaining sweeps (their kernels exit at once), while one that runs out of
    sweeps is re-solved alone in FP64 by BatchScan.finish, which costs far
    more than a few idle launches.  Then the number of lanes, by timing a
    scan of every 16th energy of every wave with those plans.  Returns
    (plans, lanes); cached in ``cache`` (JSON) when given."""
    if cache and os.path.exists(cache):
        with open(cache) as f:
            d = json.load(f)
        plans = {tuple(map(int, key.split('x'))): v for key, v in d['plans'].items()}
        return plans, int(d['lanes'])
    classes = {}
    for w in waves:
        if len(w.energies):
            classes.setdefault((w.n, w.nch), w)
    plans = {}
    for key, w in sorted(classes.items()):
        ne = len(w.energies)
        ks = sorted(set(np.linspace(0, ne - 1, min(sample, ne)).round().astype(int).tolist()))
        c = len(ks)
        bufs = {'a': torch.empty(c * w.n * w.n, dtype=torch.complex128, device='cuda'),
                'x': torch.empty(c * w.n * w.nch, dtype=torch.complex128, device='cuda'),
                'work': torch.empty(solver.batch_work_bytes('fp32', w.n, w.nch, c),
                                    dtype=torch.uint8, device='cuda')}
        base = {'trail': 0, 'subst': 0}
        timed = []
        for method in ('fp32_3m', 'fp32'):
            p = dict(base, method=method, max_refine=BATCH_REFINE_LADDER[-1])
            if not _batch_trial(w, ks, p, bufs, tol, mesh_operator, False)[0]:
                continue
            lo, hi = 0, len(BATCH_REFINE_LADDER) - 1     # hi converges
            while lo < hi:
                mid = (lo + hi) // 2
                p['max_refine'] = BATCH_REFINE_LADDER[mid]
                if _batch_trial(w, ks, p, bufs, tol, mesh_operator, False)[0]:
                    hi = mid
                else:
                    lo = mid + 1
            p['max_refine'] = BATCH_REFINE_LADDER[hi]
            ok, ms = _batch_trial(w, ks, p, bufs, tol, mesh_operator, True)
            if ok:
                timed.append((ms, p))
        if not timed:
            plans[key] = dict(base, method='fp32', max_refine=BATCH_REFINE_LADDER[-1])
            continue
        ms_best, best = min(timed, key=lambda t: t[0])
        for trail in solver.BATCH_TRAIL_TILES:
            for subst in solver.BATCH_SUBST_TILES:
                if (trail, subst) == (0, 0):
                    continue
                p = dict(best, trail=trail, subst=subst)
                ok, ms = _batch_trial(w, ks, p, bufs, tol, mesh_operator, True)
                if ok and ms < ms_best:
                    ms_best, best = ms, p
        best = dict(best, sampled_refine=best['max_refine'],
                    max_refine=min(BATCH_REFINE_LADDER[-1],
                                   max(min_refine, best['max_refine'] + margin)),
                    ms_per_system=ms_best / c)
        plans[key] = best
        if log:
            print('# batch autotune: N=%d nch=%d: %s' % (key[0], key[1], best), file=log,
                  flush=True)
        del bufs
    best_l = None
    for nl in lane_options:
        sc = BatchScan(_slice_waves(waves, 16), lanes=nl, mem_bytes=mem_bytes, tol=tol,
                       mesh_operator=mesh_operator, plans=plans)
        sc.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sc.run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if log:
            print('# batch autotune: %d lanes: %.1f ms' % (nl, dt * 1e3), file=log, flush=True)
        if best_l is None or dt < best_l[0]:
            best_l = (dt, nl)
        del sc
    if cache:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, 'w') as f:
            json.dump({'plans': {'%dx%d' % k: v for k, v in plans.items()}, 'lanes': best_l[1]}, f)
    return plans, best_l[1]


class _WaveSlice:
    """View of a wave restricted to every step-th energy (shares device data)."""

    def __init__(self, w, step):
        self.__dict__.update(w.__dict__)
        self._w = w
        self._idx = list(range(0, len(w.energies), step))
        self.energies = w.energies[self._idx]
        self.nopen = w.nopen[self._idx]
        self.qk = w.qk[self._idx]
        self.xc1 = w.xc1[self._idx]
        self.open = w.open[self._idx]
        self.co = w.co[self._idx]
        self.cop = w.cop[self._idx]
        self.nopen_dev = w.nopen_dev[self._idx]
        self.qk_host = w.qk_host[self._idx]
        self.R = torch.zeros((len(self._idx), w.nch, w.nch), dtype=torch.complex128, device='cuda')
        self.S = torch.zeros_like(self.R)
        self.res = torch.zeros(len(self._idx), dtype=torch.float64, device='cuda')
        self.status = torch.zeros(len(self._idx), dtype=torch.int32, device='cuda')

    cost = Wave.cost
    s_of = Wave.s_of


def _slice_waves(waves, step):
    return [_WaveSlice(w, step) for w in waves]


def make_waves(n_waves=64, n_energies=256, e_min=0.5, e_max=10.0, seed=0, emax=3.0):
    energies = np.linspace(e_min, e_max, n_energies)
    return [Wave(i, nc, nr, energies, emax=emax, seed=seed)
            for i, (nc, nr) in enumerate(draw_waves(n_waves, seed))]
