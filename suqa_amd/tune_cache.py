"""JSON cache of the autotuner's choices (FP64 tile bits, FP32 pass mode and
block width).

A choice is stored under a key made of the GPU's name, the workload's label,
the state size and the engine library's build signature (size and
modification time of libsuqa_hip.so), so a rebuilt library or another GPU
model tunes again.  The file is ``$SUQA_TUNE_CACHE`` when set, else
``~/.cache/suqa_amd/autotune.json``; entries are written atomically (a
temporary file renamed over the old one).  The cache only selects parameters:
results are those of the chosen configuration, which the tests cover for every
tile size and mode.
"""
import ctypes as C
import json
import os
import tempfile

from suqa_amd import capi


def cache_path():
    p = os.environ.get("SUQA_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "suqa_amd", "autotune.json")


def device_name():
    """HIP device 0: its name, else its compute capability and memory size
    (some boxes report an empty name); "unknown" where no GPU is visible."""
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = C.CDLL(name)
        except OSError:
            continue
        dev = C.c_int(0)
        if hip.hipDeviceGet(C.byref(dev), 0) != 0:
            break
        buf = C.create_string_buffer(256)
        if hip.hipDeviceGetName(buf, 256, dev) == 0 and buf.value:
            return buf.value.decode(errors="replace")
        major, minor, mem = C.c_int(0), C.c_int(0), C.c_size_t(0)
        if hip.hipDeviceComputeCapability(C.byref(major), C.byref(minor), dev) == 0 and \
                hip.hipDeviceTotalMem(C.byref(mem), dev) == 0:
            return f"hip-cc{major.value}.{minor.value}-{mem.value >> 30}GiB"
        break
    return "unknown"


def library_signature(path=capi.LIB_PATH):
    try:
        st = os.stat(path)
        return f"{st.st_size}-{int(st.st_mtime)}"
    except OSError:
        return "unbuilt"


def make_key(workload, nq, precision, device=None, lib_sig=None):
    return "|".join([device if device is not None else device_name(), workload, f"nq={nq}", precision,
                     lib_sig if lib_sig is not None else library_signature()])


def load(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def get(key, path=None):
    return load(path).get(key)


def put(key, value, path=None):
    path = path or cache_path()
    d = load(path)
    d[key] = value
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    try:
        with os.fdopen(fd, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


def tile_bits(eng, gates, candidates, workload, path=None):
    """The FP64 engine's tile bits for `workload`: the cached choice, or the
    autotuner's (suqa_hip_autotune_tile_bits, which changes the state: callers
    re-initialise after a miss) stored for next time.  Returns
    (bits, ms_per_gate_by_candidate or None, "hit" | "miss")."""
    key = make_key(workload, eng.nq, "f64-tile-bits")
    hit = get(key, path)
    if hit is not None and int(hit.get("tile_bits", -1)) in candidates:
        eng.set_tile_bits(int(hit["tile_bits"]))
        return int(hit["tile_bits"]), hit.get("ms_per_gate"), "hit"
    best, tms = eng.autotune_tile_bits(gates, candidates)
    put(key, {"tile_bits": int(best), "ms_per_gate": {str(k): round(v, 4) for k, v in tms.items()}}, path)
    return best, {str(k): round(v, 4) for k, v in tms.items()}, "miss"


def f32_config(f32, sample, workload, path=None):
    """The FP32 engine's pass mode, tile bits and block width for `workload`:
    cached, or the autotuner's (suqa_hip_f32_autotune) stored for next time.
    Returns (width, ms_per_gate or None, "hit" | "miss")."""
    key = make_key(workload, f32.nq, "f32-mode")
    hit = get(key, path)
    if hit is not None:
        f32.passes = int(hit["passes"])
        f32.tile_bits = int(hit["tile_bits"])
        return int(hit["width"]), hit.get("ms_per_gate"), "hit"
    best, tms = f32.autotune(sample)
    put(key, {"width": int(best), "passes": int(f32.passes), "tile_bits": int(f32.tile_bits),
              "ms_per_gate": {str(k): v for k, v in tms.items()} if isinstance(tms, dict) else tms}, path)
    return best, tms, "miss"
