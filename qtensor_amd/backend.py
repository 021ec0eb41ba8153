"""``HipBackend``: the bucket backend of QTensor's ``ProcessingFrameworks`` API.

Drop-in for ``NumpyBackend`` (``qtensor/ProcessingFrameworks.py:13-37``):

* ``get_sliced_buckets(buckets, data_dict, slice_dict)`` transposes every gate
  tensor to ascending variable order, applies the slice dict (``int`` drops an
  axis, ``slice`` keeps it with a new size) and uploads all tensors to the GPU
  in one host-to-device copy, as complex64 (or complex128).
* ``process_bucket(bucket, no_sum=False)`` launches one ``qb_step`` kernel: the
  broadcast product of the bucket's tensors over the union of their indices,
  summed over the bucket variable.  Returns a new ``Tensor('E<id>')`` whose
  data is a device buffer; inputs are not modified.
* ``get_result_data(result)`` copies the result to the host (numpy).

Device memory comes from PyTorch's caching allocator; kernels run on the
current torch stream.  The backend initialises the device lazily and drops
device state when pickled, so it can be handed to worker processes.
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


class HipBackend:
    def __init__(self, device=0, dtype="complex64"):
        if dtype not in ("complex64", "complex128"):
            raise ValueError("dtype must be complex64 or complex128")
        self.device_index = int(device)
        self.dtype = dtype
        self._ready = False
        self._lock = threading.Lock()
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
            self._lib = _lib.init(self.device_index)
            self.torch_dtype = torch.complex64 if self.dtype == "complex64" else torch.complex128
            self.np_dtype = np.complex64 if self.dtype == "complex64" else np.complex128
            self.qb_dtype = _lib.QB_C64 if self.dtype == "complex64" else _lib.QB_C128
            self.device = torch.device("cuda", self.device_index)
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
        torch = _torch()
        staged = []
        for bucket in buckets:
            nb = []
            for t in bucket:
                indices, data = slice_tensor_data(t, data_dict[t.data_key], slice_dict)
                nb.append((t, indices, np.ascontiguousarray(data, dtype=self.np_dtype)))
            staged.append(nb)
        flat = [d.reshape(-1) for nb in staged for (_, _, d) in nb]
        total = sum(a.size for a in flat)
        host = np.concatenate(flat) if flat else np.zeros(0, self.np_dtype)
        dev = torch.from_numpy(host).to(self.device, non_blocking=False) if total else None
        out, pos = [], 0
        for nb in staged:
            ob = []
            for t, indices, d in nb:
                view = dev[pos:pos + d.size].view(d.shape) if d.size else None
                pos += d.size
                ob.append(t.copy(indices=indices, data=view))
            out.append(ob)
        return out

    def process_bucket(self, bucket, no_sum=False):
        self._ensure()
        torch = _torch()
        if len(bucket) > _lib.QB_MAX_STEP_INPUTS:
            # multiply the head of the bucket first (no sum), then finish the step
            head = self.process_bucket(bucket[:_lib.QB_MAX_STEP_INPUTS], no_sum=True)
            return self.process_bucket([head] + list(bucket[_lib.QB_MAX_STEP_INPUTS:]), no_sum=no_sum)
        union = {}
        for t in bucket:
            for v in t.indices:
                union.setdefault(int(v), v)
        keys = sorted(union)
        U = [union[k] for k in keys]
        big = []
        for k, v in zip(keys, U):
            if v.size == 2:
                big.append(k)
            elif v.size != 1:
                raise ValueError(f"variable {v} has size {v.size}; only sizes 1 and 2 are supported")
        nb = len(big)
        pos = {k: nb - 1 - a for a, k in enumerate(big)}
        rank_u = len(big)
        if rank_u > _lib.QB_MAX_RANK:
            raise ValueError(f"bucket union rank {rank_u} exceeds {_lib.QB_MAX_RANK}")
        if U:
            tag = U[0].identity
            do_sum = not no_sum
        else:
            tag, do_sum = "f", False
        sum_flag = 1 if (do_sum and U and U[0].size == 2) else 0
        out_indices = U[1:] if do_sum else U
        desc = _lib.StepDesc()
        desc.n_inputs = len(bucket)
        desc.rank_u = rank_u
        desc.sum = sum_flag
        desc.dtype = self.qb_dtype
        ptrs = (ctypes.c_void_p * len(bucket))()
        keep = []
        for j, t in enumerate(bucket):
            m = 0
            for v in t.indices:
                k = int(v)
                if k in pos:
                    m |= 1 << pos[k]
            desc.masks[j] = m
            data = t.data
            if not isinstance(data, torch.Tensor) or not data.is_cuda:
                data = torch.as_tensor(np.asarray(data), dtype=self.torch_dtype, device=self.device)
            elif data.dtype != self.torch_dtype:
                data = data.to(self.torch_dtype)
            if not data.is_contiguous():
                data = data.contiguous()
            keep.append(data)
            ptrs[j] = data.data_ptr()
        out_shape = tuple(v.size for v in out_indices)
        out = torch.empty(out_shape, dtype=self.torch_dtype, device=self.device)
        _lib.check(self._lib.qb_step(ctypes.byref(desc), ptrs, ctypes.c_void_p(out.data_ptr()),
                                     ctypes.c_void_p(self._stream_handle())))
        return Tensor(f"E{tag}", out_indices, data=out)

    def get_result_data(self, result):
        data = result.data
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
