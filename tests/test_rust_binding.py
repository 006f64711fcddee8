"""The Rust binding a maintainer adds to the reference crate
(emerge-engine_amd/bindings/rust/hip.rs, not compiled here: the image has no
Rust toolchain) against the C-ABI it binds (include/emerge_hip.h): every
declared entry point has an `extern "C"` declaration with the same name,
argument count, argument types and return type under the C-to-Rust type
mapping, and every struct the binding mirrors has the header's fields in the
header's order with matching types."""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "emerge_hip.h")
RUST = os.path.join(REPO, "emerge-engine_amd", "bindings", "rust", "hip.rs")

STRUCTS = {
    "emh_config": "EmhConfig", "emh_material_params": "MaterialParams", "emh_material_ext": "EmhMaterialExt",
    "emh_particle": "Particle", "emh_spawn_region": "EmhSpawnRegion", "emh_step_stats": "EmhStepStats",
    "emh_snapshot": "EmhSnapshot", "emh_kernel_time": "EmhKernelTime", "emh_tuning": "EmhTuning",
    "emh_sim": "EmhSim", "emh_batch": "EmhBatch", "emh_surrogate": "EmhSurrogate",
    "emh_surrogate_train_options": "EmhTrainOptions", "emh_surrogate_counts": "EmhSurrogateCounts",
}
SCALARS = {"uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "int": "c_int", "float": "f32",
           "double": "f64", "size_t": "usize", "char": "c_char", "void": "c_void"}


def _c_type(t: str) -> str:
    """A C declaration type ('const emh_sim* const*', 'float*') as its Rust spelling."""
    t = " ".join(t.replace("*", " * ").split())
    toks = t.split()
    base = None
    level_const = False
    out_ptrs = []
    for tok in toks:
        if tok == "const":
            level_const = True
        elif tok == "*":
            out_ptrs.append(level_const)
            level_const = False
        else:
            base = tok
    base = SCALARS.get(base, STRUCTS.get(base, base))
    r = base
    for is_const in out_ptrs:
        r = ("*const " if is_const else "*mut ") + r
    return r


def _header_functions():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(
            r"\n\s*((?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**)\s*(emh_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        params = []
        a = " ".join(args.split())
        if a and a != "void":
            for p in a.split(","):
                p = p.strip()
                m = re.match(r"(.*?)([A-Za-z_][A-Za-z0-9_]*)$", p)
                params.append(_c_type(m.group(1)))
        r = _c_type(ret.strip())
        out[name] = (params, None if r == "c_void" else r)
    return out


def _rust_functions():
    src = open(RUST).read()
    block = src[src.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    out = {}
    for name, args, ret in re.findall(r"fn\s+(emh_[a-z0-9_]+)\s*\(([^)]*)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        params = []
        a = " ".join(args.split())
        if a:
            for p in a.split(","):
                p = p.strip()
                if not p:
                    continue
                params.append(" ".join(p.split(":", 1)[1].split()))
        out[name] = (params, " ".join(ret.split()) if ret else None)
    return out


def test_every_entry_point_has_a_matching_rust_declaration():
    h, r = _header_functions(), _rust_functions()
    assert len(h) >= 80
    missing = sorted(set(h) - set(r))
    assert not missing, f"no Rust extern for: {missing}"
    extra = sorted(set(r) - set(h))
    assert not extra, f"Rust declares functions the header does not: {extra}"
    for name, (params, ret) in h.items():
        rp, rr = r[name]
        assert rp == params, (name, params, rp)
        assert rr == ret, (name, ret, rr)


def _header_struct(name):
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    m = re.search(r"typedef struct %s \{(.*?)\}\s*%s;" % (name, name), src, flags=re.S)
    fields = []
    for line in m.group(1).split(";"):
        line = " ".join(line.split())
        if not line:
            continue
        first, *more = [d.strip() for d in line.split(",")]  # 'float beta1, beta2, eps'
        mm = re.match(r"(.*?)\s+([A-Za-z_][A-Za-z0-9_]*)(?:\[(\d+)\])?$", first)
        base = mm.group(1)
        for decl in [first[len(base):].strip()] + more:
            dm = re.match(r"([A-Za-z_][A-Za-z0-9_]*)(?:\[(\d+)\])?$", decl)
            t = _c_type(base)
            if dm.group(2):
                t = f"[{t}; {dm.group(2)}]"
            fields.append((dm.group(1), t))
    return fields


def _rust_struct(name):
    src = open(RUST).read()
    m = re.search(r"pub struct %s \{(.*?)\n\}" % name, src, flags=re.S)
    fields = []
    for line in m.group(1).split("\n"):
        line = line.strip().rstrip(",")
        if not line.startswith("pub "):
            continue
        fname, t = line[4:].split(":", 1)
        fields.append((fname.strip(), " ".join(t.split())))
    return fields


@pytest.mark.parametrize("c_name", ["emh_config", "emh_material_ext", "emh_spawn_region", "emh_step_stats",
                                    "emh_snapshot", "emh_kernel_time", "emh_tuning", "emh_surrogate_counts",
                                    "emh_surrogate_train_options"])
def test_mirrored_structs_match_the_header(c_name):
    assert _rust_struct(STRUCTS[c_name]) == _header_struct(c_name)
