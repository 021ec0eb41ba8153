"""``HipBackend``: the bucket backend of QTensor's ``ProcessingFrameworks`` API.

Drop-in for ``NumpyBackend`` (``qtensor/ProcessingFrameworks.py:13-37``):

* ``get_sliced_buckets(buckets, data_dict, slice_dict)`` transposes every gate
  tensor to ascending variable order, applies the slice dict (``int`` drops an
  axis, ``slice`` keeps it with a new size) and uploads all tensors to the GPU
  in one host-to-device copy, as complex64 (or complex128).
* ``process_bucket(bucket, no_sum=False)`` launches one ``qb_step`` kernel: the
  broadcast product of the bucket's tensors over the union of their indices,
  summed over the bucket variable.  Returns a new ``Tensor('E<id>')`` whose
  data is a device buffer; inputs are not modified.  A bucket of one tensor
  with nothing to add up (no sum, or a sum over a size-1 variable: the
  ket/bra wires) returns a view of that tensor's device data instead, with no
  launch.
* ``get_result_data(result)`` copies the result to the host (numpy).

Device memory comes from PyTorch's caching allocator; kernels run on the
current torch stream.  The sliced network's tensors are spans of one device
buffer (``DeviceSpan``: base buffer, element offset, shape, device address), so
no per-tensor torch view is made; step results are torch tensors.  A step's
host work is one pass over the bucket's indices, a reused per-thread
descriptor and one ``qb_step`` call.  The backend initialises the device
lazily and drops device state when pickled, so it can be handed to worker
processes; concurrent ``process_bucket`` calls from threads (the reference's
``FeynmanSimulator(pool_type='thread')``) each use their own descriptor.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from .circuit import slice_tensor_data
from .model import Tensor


def _torch():
    import torch
    return torch


class DeviceSpan:
    """Data of one sliced-network tensor: ``n`` elements at element ``offset``
    of the device buffer ``base`` (a flat torch tensor that the span keeps
    alive), laid out C-contiguous with ``shape``; ``ptr`` is its address."""
    __slots__ = ("base", "offset", "shape", "ptr")
    is_cuda = True

    def __init__(self, base, base_ptr, offset, shape, elem_bytes):
        self.base = base
        self.offset = offset
        self.shape = shape
        self.ptr = base_ptr + offset * elem_bytes

    @property
    def dtype(self):
        return self.base.dtype

    def reshaped(self, shape):
        """The same elements under another shape of the same size."""
        r = DeviceSpan.__new__(DeviceSpan)
        r.base, r.offset, r.shape, r.ptr = self.base, self.offset, shape, self.ptr
        return r

    def tensor(self):
        """A torch view of the span."""
        n = 1
        for s in self.shape:
            n *= s
        return self.base[self.offset:self.offset + n].view(self.shape)

    def __mul__(self, other):
        return self.tensor() * (other.tensor() if type(other) is DeviceSpan else other)

    def __rmul__(self, other):
        return other * self.tensor()

    def __array__(self, dtype=None, copy=None):
        a = self.tensor().cpu().numpy()
        return a if dtype is None else a.astype(dtype)


_SLAB_ELEMS = 1 << 20       # elements per result slab
_SLAB_MAX_OUT = 1 << 14     # results up to this size come from slabs


_FULL = (None, None, None)
_PERM1, _PERM2, _PERM2R = (0,), (0, 1), (1, 0)


def _bounds_key(b):
    if isinstance(b, slice):
        return (b.start, b.stop, b.step)
    return int(b)


class HipBackend:
    def __init__(self, device=0, dtype="complex64"):
        if dtype not in ("complex64", "complex128"):
            raise ValueError("dtype must be complex64 or complex128")
        self.device_index = int(device)
        self.dtype = dtype
        self._ready = False
        self._lock = threading.Lock()
        self._tls = threading.local()
        self._slice_cache = {}
        _lib.load()          # fail loudly at construction when the library is missing

    # -- pickling: no device handles cross a process boundary
    def __getstate__(self):
        return {"device_index": self.device_index, "dtype": self.dtype}

    def __setstate__(self, state):
        self.__init__(state["device_index"], state["dtype"])

    def _ensure(self):
        if self._ready:
            return
        with self._lock:
            if self._ready:
                return
            torch = _torch()
            if not torch.cuda.is_available():
                raise RuntimeError("HipBackend needs a ROCm GPU (torch.cuda.is_available() is False)")
            self._torch = torch
            self._lib = _lib.init(self.device_index)
            self.torch_dtype = torch.complex64 if self.dtype == "complex64" else torch.complex128
            self.np_dtype = np.complex64 if self.dtype == "complex64" else np.complex128
            self.qb_dtype = _lib.QB_C64 if self.dtype == "complex64" else _lib.QB_C128
            self.device = torch.device("cuda", self.device_index)
            self._elem_bytes = 8 if self.dtype == "complex64" else 16
            # qb_step with plain address arguments: the hot call passes ints
            self._qb_step = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p)(ctypes.cast(self._lib.qb_step, ctypes.c_void_p).value)
            raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
            if raw is not None:
                idx = self.device_index
                self._stream_handle = lambda: raw(idx)
            else:
                self._stream_handle = lambda: torch.cuda.current_stream(self.device).cuda_stream
            self._ready = True

    # -- BucketBackend API
    def get_sliced_buckets(self, buckets, data_dict, slice_dict):
        self._ensure()
        torch = self._torch
        staged, flat, total = [], [], 0
        cache = self._slice_cache
        bound_of = {int(v): _bounds_key(b) for v, b in slice_dict.items()}
        full = _FULL
        content = {}          # id -> values of the gate arrays seen in this call
        for bucket in buckets:
            nb = []
            for t in bucket:
                # gates repeat (same values, same index pattern, same slice), so the
                # transposed and sliced host array is computed once per pattern,
                # keyed by the gate's values rather than its data_key
                src = data_dict[t.data_key]
                tix = t.indices
                n_ax = len(tix)
                # the tensor's axes in ascending variable order
                if n_ax == 1:
                    perm = _PERM1
                elif n_ax == 2:
                    perm = _PERM2 if tix[0].identity < tix[1].identity else _PERM2R
                else:
                    ids = [v.identity for v in tix]
                    perm = tuple(sorted(range(n_ax), key=ids.__getitem__))
                bounds = tuple([bound_of.get(tix[a].identity, full) for a in perm])
                small = type(src) is np.ndarray and src.size <= 256
                hit = None
                if small:
                    val = content.get(id(src))
                    if val is None:
                        val = content[id(src)] = (src.dtype.str, src.shape, src.tobytes())
                    key = (val, perm, bounds)
                    hit = cache.get(key)
                if hit is None:
                    indices, data = slice_tensor_data(t, src, slice_dict)
                    data = np.ascontiguousarray(data, dtype=self.np_dtype)
                    sizes = tuple(v.size for v in indices)
                    kept = tuple(a for a, b in zip(perm, bounds) if not isinstance(b, int))
                    hit = (data.reshape(-1), tuple(zip(kept, sizes)), data.shape)
                    if small and len(cache) < 65536:
                        cache[key] = hit
                fdata, kept_sizes, shape = hit
                indices = tuple([tix[a] if tix[a].size == sz else tix[a].copy(size=sz) for a, sz in kept_sizes])
                nb.append((t, indices, shape, total, fdata.size))
                flat.append(fdata)
                total += fdata.size
            staged.append(nb)
        if not total:
            return [[t.copy(indices=ix, data=None) for t, ix, _, _, _ in nb] for nb in staged]
        host = np.concatenate(flat)
        dev = torch.from_numpy(host).to(self.device, non_blocking=False)
        base_ptr = dev.data_ptr()
        elem = host.itemsize
        return [[Tensor(t.name, ix, t.data_key, DeviceSpan(dev, base_ptr, off, shape, elem) if n else t.data)
                 for t, ix, shape, off, n in nb] for nb in staged]

    def _descriptor(self):
        """This thread's step descriptor and input-pointer array, with their
        addresses."""
        tls = self._tls
        try:
            return tls.step
        except AttributeError:
            desc = _lib.StepDesc()
            ptrs = (ctypes.c_void_p * _lib.QB_MAX_STEP_INPUTS)()
            tls.step = (desc, ptrs, ctypes.addressof(desc), ctypes.addressof(ptrs))
            return tls.step

    def process_bucket(self, bucket, no_sum=False):
        if not self._ready:
            self._ensure()
        if len(bucket) > _lib.QB_MAX_STEP_INPUTS:
            # multiply the head of the bucket first (no sum), then finish the step
            head = self.process_bucket(bucket[:_lib.QB_MAX_STEP_INPUTS], no_sum=True)
            return self.process_bucket([head] + list(bucket[_lib.QB_MAX_STEP_INPUTS:]), no_sum=no_sum)
        union = {}
        for t in bucket:
            for v in t.indices:
                union[v.identity] = v
        U = [union[k] for k in sorted(union)]
        # union axis a (size 2) is bit rank_u-1-a of a union offset: the last
        # size-2 axis is bit 0
        bit, rank_u = {}, 0
        for v in reversed(U):
            if v.size == 2:
                bit[v.identity] = 1 << rank_u
                rank_u += 1
            elif v.size != 1:
                raise ValueError(f"variable {v} has size {v.size}; only sizes 1 and 2 are supported")
        if rank_u > _lib.QB_MAX_RANK:
            raise ValueError(f"bucket union rank {rank_u} exceeds {_lib.QB_MAX_RANK}")
        if U:
            tag = U[0].identity
            do_sum = not no_sum
        else:
            tag, do_sum = "f", False
        sum_flag = 1 if (do_sum and U and U[0].size == 2) else 0
        out_indices = U[1:] if do_sum else U
        if sum_flag == 0 and len(bucket) == 1:
            # one tensor and nothing to add up (no sum, or a sum over a size-1
            # variable): the result holds the input's elements in the input's
            # order, so it is a view of the input's device data (no launch);
            # data are never written after they are made, so sharing is safe
            data = bucket[0].data
            if type(data) is DeviceSpan:
                return Tensor(f"E{tag}", out_indices, data=data.reshaped(tuple(v.size for v in out_indices)))
            if isinstance(data, self._torch.Tensor) and data.is_cuda and data.dtype == self.torch_dtype:
                return Tensor(f"E{tag}", out_indices, data=data.reshape([v.size for v in out_indices]))
        desc, ptrs, desc_addr, ptrs_addr = self._descriptor()
        desc.n_inputs = len(bucket)
        desc.rank_u = rank_u
        desc.sum = sum_flag
        desc.dtype = self.qb_dtype
        keep = None
        for j, t in enumerate(bucket):
            m = 0
            for v in t.indices:
                m |= bit.get(v.identity, 0)
            desc.masks[j] = m
            data = t.data
            if type(data) is DeviceSpan:
                ptrs[j] = data.ptr
                continue
            torch = self._torch
            if not isinstance(data, torch.Tensor) or not data.is_cuda:
                data = torch.as_tensor(np.asarray(data), dtype=self.torch_dtype, device=self.device)
            elif data.dtype != self.torch_dtype:
                data = data.to(self.torch_dtype)
            if not data.is_contiguous():
                data = data.contiguous()
            if keep is None:
                keep = []
            keep.append(data)         # converted inputs live until the launch is queued
            ptrs[j] = data.data_ptr()
        stream = self._stream_handle()
        n_out = 1 << (rank_u - sum_flag)
        if out_indices and n_out <= _SLAB_MAX_OUT:
            out = self._slab_span(n_out, tuple(v.size for v in out_indices), stream)
            out_ptr = out.ptr
        else:
            out = self._torch.empty([v.size for v in out_indices], dtype=self.torch_dtype, device=self.device)
            out_ptr = out.data_ptr()
        rc = self._qb_step(desc_addr, ptrs_addr, out_ptr, stream)
        if rc:
            _lib.check(rc)
        return Tensor(f"E{tag}", out_indices, data=out)

    def _slab_span(self, n, shape, stream):
        """Device memory for a small step result: the next n elements of this
        thread's current slab (a 2^20-element torch buffer that the spans keep
        alive), or of a new slab when it is full or the stream changed (a slab
        is used on one stream, the one torch's allocator orders its reuse on)."""
        tls = self._tls
        slab = getattr(tls, "slab", None)
        n_al = (n + 31) & ~31                     # 256-byte aligned in complex64
        if slab is None or slab[2] + n_al > _SLAB_ELEMS or slab[3] != stream:
            base = _torch().empty(_SLAB_ELEMS, dtype=self.torch_dtype, device=self.device)
            slab = [base, base.data_ptr(), 0, stream]
            tls.slab = slab
        off = slab[2]
        slab[2] = off + n_al
        return DeviceSpan(slab[0], slab[1], off, shape, self._elem_bytes)

    def get_result_data(self, result):
        data = result.data
        if type(data) is DeviceSpan:
            return np.asarray(data)
        if hasattr(data, "is_cuda"):
            return data.detach().cpu().numpy()
        return np.asarray(data)


class PerfBackend:
    """Per-step timing wrapper over any backend (``ProcessingFrameworks.py:152-237``).

    Times each ``process_bucket`` with device events on the current stream and
    reports bucket length, time, step flops and sizes.
    """
    Backend = HipBackend

    def __init__(self, *args, print=False, num_lines=20, **kwargs):
        self.backend = self.Backend(*args, **kwargs)
        self._print = print
        self.max_lines = num_lines
        self._profile_results = {}

    @classmethod
    def from_backend(cls, backend, *args, **kwargs):
        class CustomGeneratedBackend(cls):
            Backend = backend
        return CustomGeneratedBackend(*args, **kwargs)

    def process_bucket(self, bucket, no_sum=False):
        torch = _torch()
        indices = [tuple(t.indices) for t in bucket]
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        res = self.backend.process_bucket(bucket, no_sum=no_sum)
        end.record()
        end.synchronize()
        dt = start.elapsed_time(end) * 1e-3
        if self._print:
            print(f"PROF:: perf data process bucket time: {dt}")
        self._profile_results[str(indices)] = indices, dt
        return res

    def get_sliced_buckets(self, buckets, data_dict, slice_dict):
        return self.backend.get_sliced_buckets(buckets, data_dict, slice_dict)

    def get_result_data(self, result):
        return self.backend.get_result_data(result)

    @staticmethod
    def _perfect_bucket_flop(bucket_indices):
        union = set().union(*[set(int(v) for v in ixs) for ixs in bucket_indices])
        size_of_result = 2 ** max(len(union) - 1, 0)
        return size_of_result * (1 + len(bucket_indices))

    def gen_report(self):
        data = sorted(self._profile_results.values(), key=lambda p: p[1], reverse=True)
        lines = [str([i, ixs, t]) for i, (ixs, t) in enumerate(data[:self.max_lines])]
        rep = "\n".join(lines)
        if len(data) > self.max_lines:
            rep += f"\n ... and {len(data) - self.max_lines} lines more..."
        total = sum(d[1] for d in data)
        rep += "\n======\n"
        rep += f"Total time: {total}\nTotal bucket contractions: {len(data)}\n"
        rep += f"Mean time for contraction: {total / max(len(data), 1)}\n"
        return rep
