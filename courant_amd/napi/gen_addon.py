"""Generates courant_napi_gen.inc: one N-API wrapper per function of include/courant.h.

Parameters map by C type: `cem_sim *` and `cem_mode2d *` are small integer handles, every
other pointer is a byte offset into the arena (0 = NULL), int / long / size_t / double are
JS numbers.  Returned pointers to library-owned memory are mirrored into the arena (see
courant_napi.cpp); the sizes come from the rules in RETURNS below.  Every call that takes a
simulation writes that simulation's writable arena mirrors back before the call and
refreshes them after it.

Run: python courant_amd/napi/gen_addon.py <include/courant.h> <out.inc>
"""
import re
import sys

# returned library pointers: C expression of the mirror's length (doubles), kind id,
# index expression, writable
RETURNS = {
    "cem_port_v": ("2 * (size_t)cem_nfreq(s0)", 1, "a.I(1)", False),
    "cem_port_i": ("2 * (size_t)cem_nfreq(s0)", 2, "a.I(1)", False),
    "cem_vprobe": ("2 * (size_t)cem_nfreq(s0)", 3, "a.I(1)", False),
    "cem_iprobe": ("2 * (size_t)cem_nfreq(s0)", 4, "a.I(1)", False),
    "cem_source_spectrum": ("2 * (size_t)cem_nfreq(s0)", 5, "0", False),
    "cem_field": ("grid_points(s0)", 6, "a.I(1)", True),
    "cem_state_array": ("cem_state_len(s0, a.I(1), a.I(2), a.I(3))", 7, "a.I(1) * 64L + a.I(2) * 2 + a.I(3)", True),
    "cem_plan_array": ("plan_len(s0, a.I(1), a.I(2))", 8, "a.I(1) * 3L + a.I(2)", False),
    "cem_mode2d_eu": ("2 * (size_t)mh->ny * (mh->nz + 1)", 9, "0", False),
    "cem_mode2d_ev": ("2 * (size_t)(mh->ny + 1) * mh->nz", 10, "0", False),
    "cem_mode2d_hu": ("2 * (size_t)(mh->ny + 1) * mh->nz", 11, "0", False),
    "cem_mode2d_hv": ("2 * (size_t)mh->ny * (mh->nz + 1)", 12, "0", False),
}
STRINGS = {"cem_last_error": 20, "cem_version": 21}


def parse(header):
    src = re.sub(r"/\*.*?\*/", "", open(header).read(), flags=re.S)
    out = []
    for d in re.findall(r"CEM_API\s+([^;]+);", src):
        d = " ".join(d.split())
        m = re.match(r"(.+?)\b(cem_\w+)\s*\((.*)\)$", d)
        if not m:
            continue
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        params = [] if args in ("", "void") else [p.strip() for p in args.split(",")]
        out.append((ret, name, params))
    return out


def ptype(p):
    """C type of a parameter declaration (name dropped)."""
    t = re.sub(r"\b\w+$", "", p).strip()
    return " ".join(t.replace("*", " * ").split())


def arg_expr(t, k):
    if t in ("cem_sim *", "const cem_sim *"):
        return f"s{k}" if k == 0 else f"a.S({k})"
    if t in ("cem_mode2d *", "const cem_mode2d *"):
        return f"a.M({k})"
    if "*" in t:
        return f"static_cast<{t}>(a.Ptr({k}))"
    if t == "int":
        return f"a.I({k})"
    if t == "long":
        return f"a.L({k})"
    if t == "size_t":
        return f"a.Z({k})"
    if t == "double":
        return f"a.D({k})"
    raise ValueError(t)


def wrapper(ret, name, params):
    types = [ptype(p) for p in params]
    sim_first = bool(types) and types[0] in ("cem_sim *", "const cem_sim *")
    args = ", ".join(arg_expr(t, k) for k, t in enumerate(types))
    L = [f"napi_value js_{name}(napi_env e, napi_callback_info ci) {{", "    Call a(e, ci);"]
    if sim_first:
        L.append("    cem_sim *s0 = a.S(0);")
        L.append("    a.in->sync_in(s0);")
    if any(t in ("cem_mode2d *", "const cem_mode2d *") for t in types) and name in RETURNS:
        L.append("    const ModeH *mh = a.MH(0);")
        L.append("    if (!mh) return ret_num(e, 0);")
    call = f"{name}({args})"
    ptr_args = [str(k) for k, t in enumerate(types) if "*" in t and "cem_sim" not in t and "cem_mode2d" not in t]
    if ptr_args:
        L.append(f"    if (!a.ptrs_ok({{{', '.join(ptr_args)}}})) return ret_void(e);")
    r = ret.replace(" ", "")
    if name == "cem_destroy":
        L += ["    a.in->drop_mirrors(s0);", f"    {call};", "    a.release_sim(0);", "    return ret_void(e);"]
    elif name == "cem_mode2d_destroy":
        L += ["    a.in->drop_mirrors(a.M(0));", f"    {call};", "    a.release_mode(0);", "    return ret_void(e);"]
    elif name in STRINGS:
        L.append(f"    return ret_num(e, a.in->mirror_string({STRINGS[name]}, {call}));")
    elif name == "cem_create":
        L.append(f"    return ret_num(e, a.in->new_sim({call}));")
    elif name == "cem_mode2d_create":
        L.append(f"    return ret_num(e, a.in->new_mode({call}, a.I(0), a.I(1)));")
    elif name == "cem_cross_section_x":
        L.append(f"    cem_mode2d *m = {call};")
        L.append("    a.in->sync_out(s0);")
        L.append("    return ret_num(e, a.in->new_mode(m, a.I(3) - a.I(2), a.I(5) - a.I(4)));")
    elif name in RETURNS:
        n, kind, idx, wr = RETURNS[name]
        owner = "s0" if sim_first else "mh->m"
        L.append(f"    const double *p = {call};")
        if sim_first:
            L.append("    a.in->sync_out(s0);")
        L.append(f"    if (!p) return ret_num(e, 0);")
        L.append(f"    return ret_num(e, a.in->mirror({owner}, {kind}, {idx}, p, ({n}) * sizeof(double), "
                 f"{'true' if wr else 'false'}));")
    elif r == "void":
        L.append(f"    {call};")
        if sim_first:
            L.append("    a.in->sync_out(s0);")
        L.append("    return ret_void(e);")
    elif r in ("int", "long", "double", "size_t"):
        L.append(f"    const double r = (double)({call});")
        if sim_first:
            L.append("    a.in->sync_out(s0);")
        L.append("    return ret_num(e, r);")
    else:
        raise ValueError(f"{name}: return type {ret}")
    L.append("}")
    return "\n".join(L)


def main():
    header, out = sys.argv[1], sys.argv[2]
    fns = parse(header)
    body = ["// generated by courant_amd/napi/gen_addon.py from include/courant.h; do not edit", "namespace {", ""]
    for f in fns:
        body.append(wrapper(*f))
        body.append("")
    body.append("struct Export {\n    const char *name;\n    napi_callback fn;\n};")
    body.append("const Export kExports[] = {")
    for _, name, _ in fns:
        body.append(f'    {{"{name}", js_{name}}},')
    body.append("};")
    body.append("")
    body.append("}  // namespace")
    open(out, "w").write("\n".join(body) + "\n")
    print(f"{len(fns)} wrappers -> {out}")


if __name__ == "__main__":
    main()
