"""Host-side runtime over the C ABI: contexts, instances, batched steps.

Mirrors the reference's driver flow (ReadInput -> Allocate -> Initialize ->
time loop -> outputs, e.g. example_Dendritic_growth/kernel.cu:881-1011) as
objects: Context.create(params) allocates, Instance.upload writes the initial
fields, Context.step advances one or many instances, Instance.download reads a
field back. Every call goes through libdba_hip.so; nothing here computes.
"""
import ctypes as C

import numpy as np

from . import abi


class Instance:
    def __init__(self, ctx, iid, params):
        self.ctx = ctx
        self.id = iid
        self.params = params.copy()

    @property
    def shape(self):
        p = self.params
        return (p.nz, p.ny, p.nx) if p.nz > 1 else (p.ny, p.nx)

    @property
    def ncell(self):
        p = self.params
        return p.nx * p.ny * p.nz

    def upload(self, field, array):
        a = np.ascontiguousarray(array, dtype=np.float64).reshape(-1)
        if a.size != self.ncell:
            raise ValueError(f"field has {a.size} values, expected {self.ncell}")
        abi.check(self.ctx.lib.dba_upload(self.ctx.h, self.id, field, a.ctypes.data, a.size))

    def download(self, field):
        out = np.empty(self.ncell, dtype=np.float64)
        abi.check(self.ctx.lib.dba_download(self.ctx.h, self.id, field, out.ctypes.data, out.size))
        return out.reshape(self.shape)

    def step(self, n=1, graph_chunk=0):
        self.ctx.step([self], n, graph_chunk)

    def active(self, s=0):
        info = self.info()
        mask = np.zeros(info.ntiles, dtype=np.uint8)
        cnt = C.c_int32()
        abi.check(self.ctx.lib.dba_active_tiles(self.ctx.h, self.id, s, C.byref(cnt), mask.ctypes.data))
        return cnt.value, mask

    def worklist(self):
        info = self.info()
        buf = np.zeros(max(1, info.ntiles), dtype=np.int32)
        cnt = C.c_int32()
        abi.check(self.ctx.lib.dba_worklist(self.ctx.h, self.id, buf.ctypes.data, buf.size, C.byref(cnt)))
        return buf[: cnt.value].copy()

    def history(self):
        cap = max(1, self.params.history_cap)
        buf = np.zeros(2 * cap, dtype=np.int32)
        n = C.c_int32()
        abi.check(self.ctx.lib.dba_active_history(self.ctx.h, self.id, buf.ctypes.data, cap, C.byref(n)))
        return buf[: 2 * n.value].reshape(-1, 2).copy()

    def info(self):
        out = abi.Info()
        abi.check(self.ctx.lib.dba_instance_info(self.ctx.h, self.id, C.byref(out)))
        return out

    def check_finite(self):
        return self.ctx.lib.dba_check_finite(self.ctx.h, self.id) == abi.OK

    def field_sum(self, field):
        """Sum of the field's current values, reduced on the device in a fixed order."""
        out = C.c_double()
        abi.check(self.ctx.lib.dba_field_sum(self.ctx.h, self.id, field, C.byref(out)))
        return out.value

    def growth_front(self, xs):
        """Dendrite growth front at the given columns (at most 64)."""
        xs = np.ascontiguousarray(xs, dtype=np.int32)
        out = np.zeros(len(xs))
        abi.check(self.ctx.lib.dba_growth_front(self.ctx.h, self.id, xs.ctypes.data, len(xs), out.ctypes.data))
        return out

    def kh_vy_mode(self):
        out = C.c_double()
        abi.check(self.ctx.lib.dba_kh_vy_mode(self.ctx.h, self.id, C.byref(out)))
        return out.value

    def droplets(self, field, cap=1 << 20, noalloc=False):
        """Droplet analysis on the device: (avg, volume, sizes in the reference's
        order, mean radius). The radius follows coarsening_analysis: cbrt(3 V /
        (4 pi N)) in 3D, sqrt(V / (pi N)) in 2D (nan without droplets).
        noalloc: the entry point that works in the library's static device
        scratch (grids of at most 2^22 cells) instead of allocating."""
        avg, vol, cnt = C.c_double(), C.c_int64(), C.c_int64()
        sizes = np.zeros(max(cap, 1), dtype=np.int64)
        fn = self.ctx.lib.dba_droplets_noalloc if noalloc else self.ctx.lib.dba_droplets
        abi.check(fn(self.ctx.h, self.id, field, C.byref(avg), C.byref(vol), C.byref(cnt), sizes.ctypes.data, cap))
        n = cnt.value
        sizes = sizes[:min(n, cap)]
        if n == 0:
            r = float("nan")
        elif self.params.nz > 1:
            r = float(np.cbrt(vol.value * 0.75 / np.pi / n))
        else:
            r = (vol.value / np.pi / n) ** 0.5
        return avg.value, vol.value, sizes, r

    def destroy(self):
        if self.id is not None:
            abi.check(self.ctx.lib.dba_instance_destroy(self.ctx.h, self.id))
            self.id = None


class Context:
    """One device, one HIP stream, owned by one host thread."""

    def __init__(self, device=0, lib_path=None):
        self.lib = abi.load(lib_path)
        h = C.c_void_p()
        abi.check(self.lib.dba_init(device, C.byref(h)))
        self.h = h
        self.instances = {}

    def create(self, params):
        iid = C.c_int32()
        abi.check(self.lib.dba_instance_create(self.h, C.byref(params), C.byref(iid)))
        inst = Instance(self, iid.value, params)
        self.instances[iid.value] = inst
        return inst

    def footprint(self, params):
        b = C.c_int64()
        abi.check(self.lib.dba_footprint(C.byref(params), C.byref(b)))
        return b.value

    def device_memory(self):
        """(free, total) bytes of the context's GPU."""
        f, t = C.c_int64(), C.c_int64()
        abi.check(self.lib.dba_device_memory(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def step(self, instances, n=1, graph_chunk=0):
        ids = (C.c_int32 * len(instances))(*[i.id for i in instances])
        if graph_chunk and graph_chunk > 0:
            abi.check(self.lib.dba_step_graph(self.h, ids, len(instances), n, graph_chunk))
        else:
            abi.check(self.lib.dba_step(self.h, ids, len(instances), n))

    def sync(self):
        abi.check(self.lib.dba_sync(self.h))

    def set_single_launch(self, mode):
        """Single-launch steps: 0 (or False) forces the worklist path (step kernel
        + finalize); 1 (or True, the default) uses single-launch steps for groups of
        one instance; 2 also for batches. Results are identical in every mode."""
        abi.check(self.lib.dba_set_single_launch(self.h, int(mode)))

    def timing(self, enable):
        abi.check(self.lib.dba_timing_enable(self.h, 1 if enable else 0))

    def timing_query(self):
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_int64(), C.c_int64()
        abi.check(self.lib.dba_timing_query(self.h, C.byref(a), C.byref(na), C.byref(b), C.byref(nb)))
        return a.value, na.value, b.value, nb.value

    def marker(self, slot):
        abi.check(self.lib.dba_marker(self.h, slot))

    def elapsed_ms(self, a, b):
        ms = C.c_double()
        abi.check(self.lib.dba_marker_elapsed(self.h, a, b, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            abi.check(self.lib.dba_finalize(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
