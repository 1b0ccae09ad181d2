"""zk_gpu_amd: BN256 Fr NTT on AMD MI355X (gfx950).

Python view of libzkgpu_ntt.so (C-ABI declared in include/zkgpu_ntt.h). The
functions below mirror the reference's FFI (src/api.rs:3-13): same names,
same argument meaning, 0 on success and a nonzero code otherwise. Host
buffers are numpy uint64 arrays of shape (n, 4) (four little-endian words
per element, Montgomery form x * 2^256 mod r). Device buffers are raw
device pointers (ints), e.g. torch tensors' data_ptr().

The library must be built (python -m zk_gpu_amd.build); importing this
module without it raises: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzkgpu_ntt.so")

OK = 0
ERR_INVALID = 1
ERR_UNSUPPORTED = 2
ERR_HIP = 3
ERR_NOMEM = 4
ERR_NONCANONICAL = 5
ERR_BUDGET = 6
ERR_SHUTDOWN = 7
ERR_SCRATCH = 8

OP_FWD = 0
OP_INV = 1
OP_COSET_FWD = 2
OP_COSET_INV = 3
OPS = {"fwd": OP_FWD, "inv": OP_INV, "coset_fwd": OP_COSET_FWD, "coset_inv": OP_COSET_INV}

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_SZ = ctypes.c_size_t

# name: (restype, argtypes)
SIGNATURES = {
    "do_fft_ffi": (_U64, [_SZ]),
    "parallel_fft_api": (_U64, [_SZ, _P, _P, _P, _U64, _U64, _U64]),
    "parallel_ifft_api": (_U64, [_SZ, _P, _P, _P, _U64, _U64, _U64, _P]),
    "parallel_coeff_to_extended_part_api": (_U64, [_SZ, _P, _P, _P, _U64, _U64, _U64, _P]),
    "parallel_msm_api": (_U64, [_SZ, _P, _P, _U64, _P]),
    "zkgpu_coset_ifft_api": (_U64, [_SZ, _P, _P, _U64, _U64, _P, _P]),
    "zkgpu_ntt_device": (_I, [_I, _P, _P, _U32, _I, _P, _P, _P, _P]),
    "zkgpu_ntt_std_device": (_I, [_I, _P, _P, _U32, _I, _P]),
    "zkgpu_ntt_device_noalloc": (_I, [_I, _P, _P, _U32, _I, _P, _P, _P, _P]),
    "zkgpu_noalloc_scratch_bytes": (_U64, [_P]),
    "zkgpu_msm_device": (_I, [_I, _P, _P, _U64, _P, _P]),
    "zkgpu_msm_workspace_bytes": (_U64, [_U64, _P]),
    "zkgpu_fill_random_points": (_I, [_I, _P, _U64, _U64, _U64, _P]),
    "zkgpu_domain_constant": (_I, [_U32, _I, _P]),
    "zkgpu_prepare": (_I, [_I, _U32, _I, _P, _P, _P]),
    "zkgpu_fill_random": (_I, [_I, _P, _U64, _U64, _U64, _P]),
    "zkgpu_check_canonical": (_I, [_I, _P, _U64, _P, _P]),
    "zkgpu_plan_get": (_I, [_U32, _P, _P]),
    "zkgpu_plan_set": (_I, [_U32, _U32, _P, _P]),
    "zkgpu_plan_get_kind": (_I, [_U32, _I, _P, _P]),
    "zkgpu_plan_set_kind": (_I, [_U32, _I, _U32, _P, _P]),
    "zkgpu_batch_create": (_I, [_I, _P, _U32, _P]),
    "zkgpu_batch_launch": (_I, [_U64, _P]),
    "zkgpu_batch_info": (_I, [_U64, _P, _P, _P]),
    "zkgpu_batch_destroy": (_I, [_U64]),
    "zkgpu_ntt_batch_device": (_I, [_I, _P, _U32, _P]),
    "zkgpu_coeff_to_extended_parts_device": (_I, [_I, _P, _P, _U32, _P, _P, _U32, _P]),
    "zkgpu_coeff_to_extended_parts_api": (_U64, [_SZ, _P, _P, _P, _U64, _U64, _P, _U64]),
    "zkgpu_server_create": (_I, [_I, _U64, _U32, _U32, _P]),
    "zkgpu_server_submit": (_I, [_U64, _P, _P]),
    "zkgpu_server_poll": (_I, [_U64, _U64, ctypes.c_int32, _P]),
    "zkgpu_server_release": (_I, [_U64, _U64]),
    "zkgpu_server_stats_get": (_I, [_U64, _P]),
    "zkgpu_server_destroy": (_I, [_U64]),
    "zkgpu_table_bytes": (_U64, [_I]),
    "zkgpu_drop_tables": (_I, [_I]),
    "zkgpu_error_string": (ctypes.c_char_p, [_I]),
    "zkgpu_probe_mad_rate": (_I, [_I, _P]),
    "zkgpu_version": (_I, []),
    "zkgpu_set_validation": (_I, [_I]),
    "zkgpu_host_held_bytes": (_U64, [_I, _P]),
    "zkgpu_set_host_budget": (_U64, [_I, _U64]),
    "zkgpu_debug_fail_allocs": (_I, [_U32, _U32]),
    "zkgpu_pass_timing": (_I, [_I, _I]),
    "zkgpu_pass_times": (_I, [_I, _P, _U32, _P]),
}
OPTIONAL = {}

_lib = None


def _share_torch_hip_runtime():
    """PyTorch ships its own HIP runtime (torch/lib/libamdhip64.so, soname
    libamdhip64.so.7). The library links the system one by that soname. When
    torch is imported first, the library binds to torch's copy; loaded before
    torch, it would bring the system copy and torch a second one, and the
    library's runtime then sees no device. Loading torch's runtime file first
    (without importing torch) makes every later load, ours and torch's, use
    that one copy."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m zk_gpu_amd.build`")
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in {**SIGNATURES, **OPTIONAL}.items():
            if not hasattr(L, name):
                if name in OPTIONAL:
                    continue
                raise RuntimeError(f"{LIB_PATH} does not export {name}")
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class ZkGpuError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = lib().zkgpu_error_string(int(code)).decode()
        super().__init__(f"{what}: {msg} (code {code})" if what else f"{msg} (code {code})")


def check(code, what=""):
    if int(code) != OK:
        raise ZkGpuError(int(code), what)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return a
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], "expected a C-contiguous uint64 array"
    return a.ctypes.data_as(ctypes.c_void_p)


def _elem(a):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64).reshape(4))
    return a


# ---------------------------------------------------------------- reference FFI mirror

def do_fft_ffi(device_id=0):
    return int(lib().do_fft_ffi(device_id))


def parallel_fft_api(device_id, req, omega, twiddle, n, log_n, twiddle_size):
    """req (numpy (n,4) uint64) <- FFT_omega(req), in place (reference cuda/fft.cu:560-564)."""
    om = _elem(omega)
    return int(lib().parallel_fft_api(device_id, _ptr(req), _ptr(om), _ptr(twiddle), n, log_n, twiddle_size))


def parallel_ifft_api(device_id, req, omega, twiddle, n, log_n, twiddle_size, divisor):
    """req <- divisor * FFT_omega(req) (reference cuda/fft.cu:566-570)."""
    om, dv = _elem(omega), _elem(divisor)
    return int(lib().parallel_ifft_api(device_id, _ptr(req), _ptr(om), _ptr(twiddle), n, log_n, twiddle_size,
                                       _ptr(dv)))


def parallel_coeff_to_extended_part_api(device_id, req, omega, twiddle, n, log_n, twiddle_size, c):
    """req[i] *= c^i, then req <- FFT_omega(req) (reference cuda/fft.cu:572-576)."""
    om, cc = _elem(omega), _elem(c)
    return int(lib().parallel_coeff_to_extended_part_api(device_id, _ptr(req), _ptr(om), _ptr(twiddle), n, log_n,
                                                         twiddle_size, _ptr(cc)))


def parallel_msm_api(device_id, data, bases, length, acc):
    """acc (numpy (3, 4) uint64: x, y, z) <- sum_i data[i] * bases[i] over G1;
    data (length, 4) uint64 Fr Montgomery words, bases (length, 8) uint64 (x
    then y, Fq Montgomery words) (reference cuda/fft.cu:450-551)."""
    return int(lib().parallel_msm_api(device_id, _ptr(data), _ptr(bases), length, _ptr(acc)))


def fill_random_points(d_out, count, seed, start=0, stream=None, device_id=0):
    """Synthetic affine bases k_i G (64 bytes each) at device pointer d_out."""
    check(lib().zkgpu_fill_random_points(device_id, d_out, count, seed, start, stream), "zkgpu_fill_random_points")


def msm_device(d_scalars, d_bases, length, stream=None, device_id=0):
    """MSM on device pointers; returns the (3, 4) uint64 result (x, y, z)."""
    acc = np.zeros((3, 4), dtype=np.uint64)
    check(lib().zkgpu_msm_device(device_id, d_scalars, d_bases, length, _ptr(acc), stream), "zkgpu_msm_device")
    return acc


def zkgpu_coset_ifft_api(device_id, req, omega, n, log_n, divisor, c):
    om, dv, cc = _elem(omega), _elem(divisor), _elem(c)
    return int(lib().zkgpu_coset_ifft_api(device_id, _ptr(req), _ptr(om), n, log_n, _ptr(dv), _ptr(cc)))


# ---------------------------------------------------------------- device-resident API

def domain_constant(log_n, which):
    """which: 0 omega, 1 omega^-1, 2 n^-1, 3 g=7, 4 g^-1 (Montgomery words)."""
    out = np.zeros(4, dtype=np.uint64)
    check(lib().zkgpu_domain_constant(log_n, which, _ptr(out)), "zkgpu_domain_constant")
    return out


def ntt_device(d_in, d_out, log_n, op, omega, divisor=None, c=None, stream=None, device_id=0):
    om, dv, cc = _elem(omega), _elem(divisor), _elem(c)
    check(lib().zkgpu_ntt_device(device_id, d_in, d_out, log_n, op, _ptr(om), _ptr(dv), _ptr(cc), stream),
          "zkgpu_ntt_device")


def ntt_std_device(d_in, d_out, log_n, op, stream=None, device_id=0):
    check(lib().zkgpu_ntt_std_device(device_id, d_in, d_out, log_n, op, stream), "zkgpu_ntt_std_device")


def prepare_std(log_n, op, device_id=0):
    inverse = op in (OP_INV, OP_COSET_INV)
    om = domain_constant(log_n, 1 if inverse else 0)
    dv = domain_constant(log_n, 2)
    cc = domain_constant(log_n, 4 if op == OP_COSET_INV else 3)
    check(lib().zkgpu_prepare(device_id, log_n, op, _ptr(om), _ptr(dv), _ptr(cc)), "zkgpu_prepare")


def fill_random(d_out, count, seed, start=0, stream=None, device_id=0):
    check(lib().zkgpu_fill_random(device_id, d_out, count, seed, start, stream), "zkgpu_fill_random")


PLAN_SINGLE = 0
PLAN_BATCH = 1


def plan_get(log_n, kind=PLAN_SINGLE):
    d = np.zeros(3, dtype=np.uint32)
    c = np.zeros(3, dtype=np.uint32)
    npass = lib().zkgpu_plan_get_kind(log_n, kind, d.ctypes.data_as(ctypes.c_void_p),
                                      c.ctypes.data_as(ctypes.c_void_p))
    return int(npass), [int(x) for x in d[:npass]], [int(x) for x in c[:npass]]


def plan_set(log_n, digits, log_cols, kind=PLAN_SINGLE):
    d = np.array(list(digits) + [0] * (3 - len(digits)), dtype=np.uint32)
    c = np.array(list(log_cols) + [0] * (3 - len(log_cols)), dtype=np.uint32)
    check(lib().zkgpu_plan_set_kind(log_n, kind, len(digits), d.ctypes.data_as(ctypes.c_void_p),
                                    c.ctypes.data_as(ctypes.c_void_p)), "zkgpu_plan_set")


def plan_reset(log_n, kind=PLAN_SINGLE):
    check(lib().zkgpu_plan_set_kind(log_n, kind, 0, None, None), "zkgpu_plan_set")


# ---------------------------------------------------------------- batches

JOB_STD_DOMAIN = 1


class Job(ctypes.Structure):
    """zkgpu_job (include/zkgpu_ntt.h)."""
    _fields_ = [("d_in", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("log_n", ctypes.c_uint32),
                ("op", ctypes.c_int32), ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("omega", ctypes.c_uint64 * 4), ("divisor", ctypes.c_uint64 * 4), ("coset", ctypes.c_uint64 * 4)]


JOB_HOST_BUFFERS = 2


def make_job(d_in, d_out, log_n, op, omega=None, divisor=None, c=None, host=False):
    """A job on device pointers (host=True: host buffers, server jobs only,
    pointers as ints or uint64 numpy arrays); without constants it uses the
    standard domain (g = 7)."""
    j = Job()
    if host:
        d_in = d_in if isinstance(d_in, int) else d_in.ctypes.data
        d_out = d_out if isinstance(d_out, int) else d_out.ctypes.data
    j.d_in, j.d_out, j.log_n, j.op = d_in, d_out, log_n, op
    j.flags = JOB_HOST_BUFFERS if host else 0
    if omega is None:
        j.flags |= JOB_STD_DOMAIN
    else:
        for name, v in (("omega", omega), ("divisor", divisor), ("coset", c)):
            if v is not None:
                getattr(j, name)[:] = [int(x) for x in _elem(v)]
    return j


class Batch:
    """Many device-resident transforms as one captured launch graph."""

    def __init__(self, jobs, device_id=0):
        arr = (Job * len(jobs))(*jobs)
        self._jobs = arr
        h = ctypes.c_uint64(0)
        check(lib().zkgpu_batch_create(device_id, arr, len(jobs), ctypes.byref(h)), "zkgpu_batch_create")
        self.handle = h.value

    def launch(self, stream=None):
        check(lib().zkgpu_batch_launch(self.handle, stream), "zkgpu_batch_launch")

    def info(self):
        lv, tl, el = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint64(0)
        check(lib().zkgpu_batch_info(self.handle, ctypes.byref(lv), ctypes.byref(tl), ctypes.byref(el)))
        return {"levels": lv.value, "tiles": tl.value, "elements": el.value}

    def close(self):
        if self.handle:
            lib().zkgpu_batch_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coeff_to_extended_parts_device(d_in, d_outs, log_n, omega, cosets, stream=None, device_id=0):
    """part j = FFT_omega(a[i] * cosets[j]^i) on device pointers, one batched run."""
    outs = (ctypes.c_void_p * len(d_outs))(*d_outs)
    cs = np.ascontiguousarray(np.asarray(cosets, dtype=np.uint64).reshape(len(d_outs), 4))
    check(lib().zkgpu_coeff_to_extended_parts_device(device_id, d_in, outs, log_n, _ptr(_elem(omega)), _ptr(cs),
                                                     len(d_outs), stream), "zkgpu_coeff_to_extended_parts_device")


def coeff_to_extended_parts_api(device_id, coeffs, omega, n, log_n, cosets, out=None):
    """Host version: returns a (count, n, 4) uint64 array of parts (written
    into `out` when given)."""
    cs = np.ascontiguousarray(np.asarray(cosets, dtype=np.uint64).reshape(-1, 4))
    parts = np.empty((cs.shape[0], n, 4), dtype=np.uint64) if out is None else out
    assert parts.shape == (cs.shape[0], n, 4) and parts.dtype == np.uint64 and parts.flags["C_CONTIGUOUS"]
    check(lib().zkgpu_coeff_to_extended_parts_api(device_id, _ptr(np.ascontiguousarray(coeffs)), _ptr(parts),
                                                  _ptr(_elem(omega)), n, log_n, _ptr(cs), cs.shape[0]),
          "zkgpu_coeff_to_extended_parts_api")
    return parts


# ---------------------------------------------------------------- serving

class ServerResult(ctypes.Structure):
    """zkgpu_server_result (include/zkgpu_ntt.h)."""
    _fields_ = [("status", ctypes.c_int32), ("done", ctypes.c_int32), ("d_out", ctypes.c_void_p),
                ("t_submit_ns", ctypes.c_uint64), ("t_admit_ns", ctypes.c_uint64),
                ("t_start_ns", ctypes.c_uint64), ("t_done_ns", ctypes.c_uint64),
                ("iterations", ctypes.c_uint32), ("iter_jobs_max", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    """zkgpu_server_stats (include/zkgpu_ntt.h)."""
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "submitted", "completed", "failed", "pending", "active", "iterations", "budget_bytes",
        "in_use_bytes", "peak_in_use_bytes", "pool_bytes", "elements_done")]


class Server:
    """Continuous-batching NTT server on one device (jobs as make_job();
    d_out = None asks for a server-owned result buffer)."""

    def __init__(self, budget_bytes=0, max_active=0, device_id=0):
        h = ctypes.c_uint64(0)
        check(lib().zkgpu_server_create(device_id, budget_bytes, max_active, 0, ctypes.byref(h)),
              "zkgpu_server_create")
        self.handle = h.value

    def submit(self, job):
        t = ctypes.c_uint64(0)
        check(lib().zkgpu_server_submit(self.handle, ctypes.byref(job), ctypes.byref(t)), "zkgpu_server_submit")
        return t.value

    def poll(self, ticket, wait_ms=0):
        r = ServerResult()
        check(lib().zkgpu_server_poll(self.handle, ticket, wait_ms, ctypes.byref(r)), "zkgpu_server_poll")
        return r

    def wait(self, ticket):
        r = self.poll(ticket, -1)
        check(r.status, "server job")
        return r

    def release(self, ticket):
        check(lib().zkgpu_server_release(self.handle, ticket), "zkgpu_server_release")

    def stats(self):
        st = ServerStats()
        check(lib().zkgpu_server_stats_get(self.handle, ctypes.byref(st)), "zkgpu_server_stats_get")
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().zkgpu_server_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def probe_mad_rate(device_id=0):
    """Measured v_mad_u64_u32 throughput of the device (multiply-adds/s)."""
    v = ctypes.c_double(0)
    check(lib().zkgpu_probe_mad_rate(device_id, ctypes.byref(v)), "zkgpu_probe_mad_rate")
    return v.value


class KernelTime(ctypes.Structure):
    """zkgpu_kernel_time (include/zkgpu_ntt.h)."""
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("total_ms", ctypes.c_double)]


def pass_timing(enable, device_id=0):
    """Turn per-pass kernel timing (events tied to each pass dispatch) on or off."""
    check(lib().zkgpu_pass_timing(device_id, 1 if enable else 0), "zkgpu_pass_timing")


def pass_times(device_id=0):
    """{kernel name: (launches, total ms)} of the passes timed since the last read."""
    buf = (KernelTime * 32)()
    cnt = ctypes.c_uint32(0)
    check(lib().zkgpu_pass_times(device_id, buf, 32, ctypes.byref(cnt)), "zkgpu_pass_times")
    return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(min(cnt.value, 32))}


def version():
    return int(lib().zkgpu_version())
