"""The Julia binding (geophyinv.jl_amd/julia/GPI.jl) against the C header it binds
(include/gpi.h): enum ids, face bits and the field order of every struct it mirrors.
A Julia toolchain is not available, so the binding's text is checked, not run."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = open(os.path.join(ROOT, "include", "gpi.h")).read()
JL = open(os.path.join(ROOT, "geophyinv.jl_amd", "julia", "GPI.jl")).read()

STRUCTS = {"Grid": "gpi_grid", "CArray": "gpi_array", "CPml": "gpi_pml", "Points": "gpi_points",
           "Shot": "gpi_shot", "AdjShot": "gpi_adjoint_shot"}


def c_enum(name):
    body = re.search(r"enum %s \{(.*?)\};" % name, HEADER, re.S).group(1)
    ids, k = {}, 0
    for tok in re.findall(r"(GPI_\w+)(?:\s*=\s*(\d+))?", body):
        if tok[1]:
            k = int(tok[1])
        ids[tok[0]] = k
        k += 1
    return ids


def jl_dict(name):
    body = re.search(r"const %s = Dict\((.*?)\)\n" % name, JL, re.S).group(1)
    return {k: int(v) for k, v in re.findall(r":(\w+)\s*=>\s*(\d+)", body)}


def c_struct_fields(name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), HEADER, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            m = re.search(r"(\w+)\s*(\[\d+\])?\s*$", part.strip())
            names.append(m.group(1))
    return names


def jl_struct_fields(name):
    body = re.search(r"struct %s\n(.*?)\nend" % name, JL, re.S).group(1)
    return re.findall(r"^\s*(\w+)::", body, re.M)


def test_field_ids_match_the_header():
    c = c_enum("gpi_field")
    jl = jl_dict("FIELD")
    for name, v in jl.items():
        assert c["GPI_" + name.upper()] == v, name
    assert len(jl) == c["GPI_NFIELDS"]


def test_coefficient_ids_match_the_header():
    c = c_enum("gpi_coef")
    jl = jl_dict("COEF")
    assert len(jl) == c["GPI_NCOEF"]
    for name, v in jl.items():
        assert c["GPI_C_" + name.upper()] == v, name
    # every name of MediumParameters (src/fdtd/medium.jl:81-95) has an id
    for names in (["invK", "rho", "dtinvavxirho", "dtinvavyirho", "dtinvavzirho", "dtK"],
                  ["invlambda", "invmu", "rho", "dtinvavxirho", "dtinvavyirho", "dtinvavzirho", "dtavmu",
                   "dtavxzimu", "dtavxyimu", "dtavyzimu", "dtM", "dtlambda"]):
        assert all(n in jl for n in names)


def test_face_bits_match_the_header():
    bits = jl_dict("FACE_BIT")
    for face, b in bits.items():
        axis = {"z": 0, "y": 1, "x": 2}[face[0]]
        side = 0 if face.endswith("min") else 1
        assert b == 2 * axis + side   # GPI_FACE(axis, side)


@pytest.mark.parametrize("jl,c", sorted(STRUCTS.items()))
def test_struct_layouts_match_the_header(jl, c):
    assert jl_struct_fields(jl) == c_struct_fields(c)


def test_binding_passes_mod_and_dmod():
    """gpi_model_create receives every pac.dmod and every pac.mod array (invK or invlambda
    and invmu, and rho), which is the set the elastic passes need."""
    body = re.search(r"function coef_arrays\(pac\)(.*?)\nend", JL, re.S).group(1)
    assert "names(pac.dmod)[1]" in body and "names(pac.mod)[1]" in body


def test_binding_has_no_placeholder_bodies():
    # splats (f(x...)) are code; a line that is only an ellipsis, or an ellipsis comment, is not
    assert not re.search(r"^\s*(#\s*)?\.\.\.", JL, re.M)


def test_reverse_rule_calls_functions_of_the_binding():
    """GPIChainRules.jl (the ChainRulesCore rule of lossvalue) calls only functions that
    GPI.jl defines, and its pullback is the adjoint half of gradient! (func_grad.jl:21-45):
    adjoint sources, time reversal, the :adjoint run, the exp(m) * ref_mod chain rule."""
    rules = open(os.path.join(ROOT, "geophyinv.jl_amd", "julia", "GPIChainRules.jl")).read()
    assert not re.search(r"^\s*(#\s*)?\.\.\.", rules, re.M)
    defined = set(re.findall(r"^function (\w+!?)\(", JL, re.M))
    used = set(re.findall(r"\bGPI\.(\w+!?)\(", rules))
    assert used and used <= defined, used - defined
    body = re.search(r"function backprop!\((.*?)\nend", JL, re.S).group(1)
    for step in ("GP.gradient!(pac.srcwav[2]", "reverse!(pac.srcwav[2])", "[-1, 1]", ":adjoint",
                 "exp(x1) * r"):
        assert step in body, step


def c_prototypes():
    """Entry point name -> number of parameters, from the header's prototypes."""
    out = {}
    text = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    for m in re.finditer(r"\b(?:int|void|const char\s*\*|uint32_t)\s+(gpi_\w+)\s*\((.*?)\);", text, re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def jl_ccalls():
    """(name, number of argument types) of every ccall in the binding."""
    calls = []
    for m in re.finditer(r"ccall\(\(:(gpi_\w+), lib\),\s*\w+,\s*\(", JL):
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(JL[i], 0)
            i += 1
        types = JL[m.end():i - 1]
        level, n = 0, 1 if types.strip() else 0
        for ch in types:
            level += {"{": 1, "}": -1, "(": 1, ")": -1}.get(ch, 0)
            if ch == "," and level == 0:
                n += 1
        if types.strip().endswith(","):
            n -= 1
        calls.append((m.group(1), n))
    return calls


def test_every_ccall_matches_a_header_prototype():
    """Each entry point the binding calls is declared in the header with as many parameters
    as the ccall passes types (the FWI outer loop's included: conv_lsmr!, interp_apply!,
    the L-BFGS handle)."""
    protos = c_prototypes()
    calls = jl_ccalls()
    names = {n for n, _ in calls}
    assert {"gpi_forward", "gpi_adjoint", "gpi_conv_lsmr", "gpi_interp_apply", "gpi_lbfgs_direction"} <= names
    for name, n in calls:
        assert name in protos, name
        assert protos[name] == n, (name, protos[name], n)
