"""Neural radiance cache for the NEE volume path tracer.

An MLP Q(x, w) -> RGB (96 Fourier/direction features, 4 hidden ReLU layers of
64, BF16 weights and activations with FP32 accumulation on MFMA) learns, online
from the path tracer's own samples, the radiance a NEE path collects from a real
collision at x reached along w.  Rendering stops each path at its
(terminate_after + 1)-th collision and adds the path throughput times Q there;
the plain NEE integrator (DeltatrackingNEEPathIntegrator) stays the reference
whose error the cached render is measured against.  No reference counterpart:
rendervous has no radiance cache.  Kernels: vr_nrc_* in include/vr_api.h.
"""
from __future__ import annotations

import math
import typing

import torch

from . import _internal
from . import _lib as L
from . import maps as M

# layer shapes (outputs, inputs) in parameter order; the last layer has 16
# outputs of which the first 3 are RGB
LAYERS = ((64, 96), (64, 64), (64, 64), (64, 64), (16, 64))


def _f3p(v) -> "ctypes.Array":
    return L.f3(v)


class NeuralRadianceCache:
    """Weights, Adam state and the training / query buffers of one cache."""

    MAX_QUERIES = 1 << 23  # rows of the render's query buffer (x, w, W/spp, pixel: 40 B each)

    def __init__(self, device: typing.Optional[torch.device] = None, *, seed: int = 0, lr: float = 1e-2,
                 betas=(0.9, 0.99), eps: float = 1e-8, terminate_after: int = 2):
        device = torch.device("cuda", 0) if device is None else torch.device(device)
        g = torch.Generator().manual_seed(seed)
        chunks = []
        for i, (n, k) in enumerate(LAYERS):
            bound = math.sqrt(6.0 / k) * (0.1 if i == len(LAYERS) - 1 else 1.0)  # He-uniform for ReLU
            w = (torch.rand(n, k, generator=g) * 2.0 - 1.0) * bound
            if i == len(LAYERS) - 1:
                w[3:] = 0.0
            chunks.append(w.reshape(-1))
        self.device = device
        self.params = torch.cat(chunks).to(device)
        assert self.params.numel() == L.NRC_PARAMS
        self.weights = self.params.to(torch.bfloat16)  # round to nearest even, as the kernels
        # a training step accumulates into NRC_GRAD_COPIES copies; the first
        # holds the step's gradient afterwards (self.grad views it)
        self.grad_copies = torch.zeros(L.NRC_GRAD_COPIES * L.NRC_PARAMS, device=device)
        self.grad = self.grad_copies[:L.NRC_PARAMS]
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        # a training step's loss buffer; its mean loss lands in the first slot
        self._loss_slots = torch.zeros(L.NRC_LOSS_SLOTS, device=device)
        self.loss = self._loss_slots[:1]
        self.lr, self.betas, self.eps = lr, betas, eps
        self.terminate_after = int(terminate_after)
        self.steps = 0
        self._records = None
        self._records_cap = 0
        self._last_traces = 1
        self._scratch = None
        self._count = torch.zeros(1, dtype=torch.int32, device=device)
        self._queries = None
        self._qcount = torch.zeros(1, dtype=torch.int32, device=device)
        self._qtotal = torch.zeros(1, dtype=torch.int64, device=device)

    # --------------------------------------------------------------- helpers
    @staticmethod
    def _lower(field: M.MapBase):
        kind, lw = field._lower_map()
        if kind != L.VR_MAP_NEE_DT:
            raise NotImplementedError("the radiance cache serves the NEE path integrator (DeltatrackingNEEPathIntegrator)")
        return lw

    def _stream(self):
        return L.stream_handle(self.device)

    # -------------------------------------------------------------- training
    def train(self, field: M.MapBase, camera: M.PerspectiveCameraSensor, n_paths: int = 8192, max_vertices: int = 16,
              steps: int = 1, batch: typing.Optional[int] = None,
              lr: typing.Union[None, float, typing.Sequence[float]] = None, traces: int = 1,
              seeds: typing.Optional[typing.Sequence[typing.Sequence[int]]] = None) -> torch.Tensor:
        """Traces `traces` sets of n_paths plain NEE paths from random sensors
        and, for each set in turn, runs `steps` Adam steps on its collision
        vertices, each on the next `batch` records (all records when batch is
        None); returns the last mean loss (device).  The paths do not read the
        cache, so the sets are traced in one launch (up to 16), with the result
        of `traces` calls of one set each.  `seeds` gives each set's dispatch
        seeds (drawn with get_seeds() when None).  `lr` overrides the cache's
        learning rate: one value, or one per step of all sets (a schedule)."""
        traces = int(traces)
        if not 1 <= traces <= 16:
            raise ValueError("traces must be in [1, 16]")
        lw = self._lower(field)
        cam, keep = camera._camera(None)
        cap = n_paths * max_vertices
        if self._records is None or self._records.shape[0] < traces * cap or self._records_cap != cap:
            self._records = torch.empty(traces * cap, 9, device=self.device)
            self._scratch = torch.empty(traces * cap * 12, device=self.device)
            self._records_cap = cap
        if self._count.numel() < traces:
            self._count = torch.zeros(traces, dtype=torch.int32, device=self.device)
        if seeds is None:
            seeds = [_internal.get_seeds() for _ in range(traces)]
        if len(seeds) != traces:
            raise ValueError(f"{len(seeds)} seed tuples for {traces} traces")
        self._last_traces = traces
        launches = (L.VrLaunch * traces)(*[L.launch(sd, bw_mode=L.VR_BW_CONSISTENT) for sd in seeds])
        lib = L.load()
        L.check(lib.vr_nrc_paths_traces(lw.medium, cam, launches, traces, int(n_paths), int(max_vertices),
                                        self._scratch.data_ptr(), self._records.data_ptr(), int(cap),
                                        self._count.data_ptr(), self._stream()), "vr_nrc_paths_traces")
        bmin, bmax = _f3p(lw.medium.box_min), _f3p(lw.medium.box_max)
        size = cap if batch is None else int(batch)
        n_steps = traces * steps
        if lr is None or isinstance(lr, (int, float)):
            rates = [self.lr if lr is None else float(lr)] * n_steps
        else:
            rates = [float(r) for r in lr]
            if len(rates) != n_steps:
                raise ValueError(f"lr schedule has {len(rates)} entries for {n_steps} steps")
        row_bytes = 9 * self._records.element_size()
        for t in range(traces):
            records = self._records.data_ptr() + t * cap * row_bytes
            count = self._count.data_ptr() + t * self._count.element_size()
            for k in range(steps):
                self.steps += 1
                first = 0 if batch is None else k * size  # wrapped at the written count by the kernel
                L.check(lib.vr_nrc_train_step(records, count, int(first), int(size), bmin, bmax,
                                              self.params.data_ptr(), self.weights.data_ptr(),
                                              self.grad_copies.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                              self.steps, rates[t * steps + k], self.betas[0], self.betas[1], self.eps,
                                              self.loss.data_ptr(), self._stream()), "vr_nrc_train_step")
        return self.loss

    def samples_last(self) -> int:
        """Training records written by the last train(), over its traces (host sync)."""
        return int(self._count[:self._last_traces].to(torch.int64).sum().item())

    # ------------------------------------------------------------- rendering
    def render(self, field: M.MapBase, camera: M.PerspectiveCameraSensor, spp: int = 1) -> torch.Tensor:
        """Capture of the NEE radiance with paths ended into the cache.

        A sample issues at most one query, so the sensors are rendered in
        ranges of at most MAX_QUERIES // spp sensors: the query buffer never
        exceeds MAX_QUERIES rows (40 B each) and is allocated once.  The ranges
        use the same dispatch seeds and per-sample streams as one launch."""
        lw = self._lower(field)
        cam, keep = camera._camera(None)
        total = camera.num_sensors()
        spp = max(1, int(spp))
        per = max(1, min(total, self.MAX_QUERIES // spp))
        need = per * spp
        if self._queries is None or self._queries.shape[0] < need:
            self._queries = torch.empty(max(need, min(self.MAX_QUERIES, total * spp)), 10, device=self.device)
        out = torch.empty(*camera.index_shape, 3, device=self.device)
        seeds = _internal.get_seeds()
        self._qtotal.zero_()
        lib = L.load()
        with M._ENGINE_LOCK:
            for start in range(0, total, per):
                launch = L.launch(seeds, samples=spp, bw_mode=L.VR_BW_CONSISTENT, start_index=start,
                                  count=min(per, total - start))
                L.check(lib.vr_nrc_render(lw.medium, cam, launch, self.terminate_after, self.weights.data_ptr(),
                                          out.data_ptr(), self._queries.data_ptr(), int(need), self._qcount.data_ptr(),
                                          self._stream()), "vr_nrc_render")
                self._qtotal += self._qcount.to(torch.int64)
        return out

    def queries_last(self) -> int:
        """Cache queries issued by the last render() (host sync)."""
        return int(self._qtotal.item())
