"""Write tests/golden/abi_closure.json: the functions a reference host that
links libopenstaple_amd still needs from the reference files it drops.

The reference program (src/Makefile.am:46-184, main_SOURCES) is linked with
the library after removing the files and single definitions listed below
(the same list as the integration guide).  For every non-static function
defined in a removed file (or removed definition) this script records which
kept source files call it.  It also records the single-precision twins that
the kept code calls: the reference's generator writes sp_<file> from <file>
with every function name suffixed "_f" (double_to_single_transformer.py:14-31),
so a kept file that the generator converts calls `name_f` wherever the double
file calls `name`, and hand-written kept files may call `name_f` directly.

Run from the repository root, in a checkout with the reference tree at
/root/reference:  python tests/golden/make_abi_closure.py
The output is data (names of functions and files), not reference source.
"""
from __future__ import annotations

import json
import pathlib
import re
import sys

REF = pathlib.Path("/root/reference/src")
OUT = pathlib.Path(__file__).resolve().parent / "abi_closure.json"

# Reference files the library replaces as a whole
DROPPED_FILES = [
    "OpenAcc/fermion_matrix.c",
    "OpenAcc/fermionic_utilities.c",
    "OpenAcc/inverter_full.c",
    "OpenAcc/inverter_mixedp.c",
    "OpenAcc/inverter_multishift_full.c",
    "OpenAcc/inverter_wrappers.c",
    "OpenAcc/stouting.c",
    "OpenAcc/fermion_force.c",
    "OpenAcc/fermion_force_utilities.c",
    "OpenAcc/find_min_max.c",
    "RationalApprox/rationalapprox.c",
    "OpenAcc/ipdot_gauge.c",
    "OpenAcc/rectangles.c",
    "OpenAcc/md_integrator.c",
    "OpenAcc/md_parameters.c",
    "OpenAcc/update_versatile.c",
    "OpenAcc/io.c",
    "Meas/baryon_number_utilities.c",
    "OpenAcc/field_times_fermion_matrix.c",
    # the hand-written single-precision fermion force: only the generated
    # twin of md_integrator.c calls it
    "OpenAcc/sp_fermion_force.c",
]
# Kept files whose generated single-precision twin is dropped as well: only
# the dropped single-precision MD calls into them
DROPPED_TWINS_OF_KEPT = ["OpenAcc/topological_action.c", "OpenAcc/topological_force.c"]
# Single definitions the library replaces inside files that stay
DROPPED_DEFS = {
    "OpenAcc/su3_measurements.c": ["calc_plaquette_soloopenacc", "calc_momenta_action", "invert_momenta"],
    "OpenAcc/su3_utilities.c": ["mom_sum_mult", "mom_exp_times_conf_soloopenacc"],
    "DbgTools/dbgtools.c": ["print_vec3_soa", "read_vec3_soa"],
    "Meas/magnetic_susceptibility_utilities.c": ["idphase_dbz"],
}

KEYWORDS = {"if", "for", "while", "switch", "return", "sizeof", "else", "do"}


def strip_c(text: str) -> str:
    """Remove comments, string/char literals and preprocessor lines."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r'"(\\.|[^"\\\n])*"', '""', text)
    text = re.sub(r"'(\\.|[^'\\\n])*'", "''", text)
    out = []
    cont = False
    for line in text.split("\n"):
        if cont or line.lstrip().startswith("#"):
            cont = line.rstrip().endswith("\\")
            out.append("")
            continue
        out.append(line)
    return "\n".join(out)


def _variable(decl: str) -> tuple[str, bool] | None:
    """Name and staticness of a top-level variable definition (the text
    before its initialiser), None for prototypes, typedefs and externs."""
    d = decl.strip()
    if not d or re.match(r"(typedef|extern)\b", d) or re.search(r"\b(struct|union|enum)\s*\w*\s*$", d):
        return None
    st = bool(re.search(r"\bstatic\b", d))
    if "(" in d:
        m = re.search(r"\(\s*\*\s*([A-Za-z_]\w*)", d)  # pointer to function (or array of them)
        return (m.group(1), st) if m else None
    m = re.search(r"([A-Za-z_]\w*)\s*(\[[^\]]*\]\s*)*$", d)
    if not m or not re.search(r"[A-Za-z_]\w*\s", d[: m.start()] + " "):
        return None
    return (m.group(1), st)


def definitions(path: pathlib.Path, variables: bool = False) -> dict[str, bool]:
    """Top-level function definitions of a C file (and, with `variables`,
    global variable definitions): name -> is_static.  Every branch of a
    preprocessor conditional is scanned."""
    s = strip_c(path.read_text(errors="replace"))
    defs: dict[str, bool] = {}

    def add(name: str, st: bool) -> None:
        defs[name] = defs.get(name, False) or st

    depth = 0
    i = 0
    stmt_start = 0
    n = len(s)
    while i < n:
        ch = s[i]
        if ch == "{":
            if depth == 0:
                head = s[stmt_start:i]
                m = re.search(r"([A-Za-z_]\w*)\s*\(([^()]|\([^()]*\))*\)\s*$", head)
                if m and m.group(1) not in KEYWORDS and "=" not in head[: m.start()]:
                    name = m.group(1)
                    before = head[: m.start()]
                    if re.search(r"[A-Za-z_]\w*", before):  # has a return type
                        add(name, bool(re.search(r"\bstatic\b", before)))
                elif variables and "=" in head:
                    v = _variable(head[: head.index("=")])
                    if v:
                        add(*v)
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                stmt_start = i + 1
        elif ch == ";" and depth == 0:
            stmt = s[stmt_start:i]
            if variables and "{" not in stmt:
                v = _variable(stmt.split("=")[0])
                if v:
                    add(*v)
            stmt_start = i + 1
        i += 1
    return defs


def body_words(text: str) -> dict[str, int]:
    """Whole-word identifier counts inside braces (function bodies and
    initialisers): calls and function-pointer references, not prototypes or
    the definition's own head.  `text` is already stripped."""
    counts: dict[str, int] = {}
    depth = 0
    for m in re.finditer(r"[{}]|[A-Za-z_]\w*", text):
        tok = m.group(0)
        if tok == "{":
            depth += 1
        elif tok == "}":
            depth -= 1
        elif depth > 0:
            counts[tok] = counts.get(tok, 0) + 1
    return counts


def local_includes(path: pathlib.Path, seen: set[pathlib.Path]) -> list[pathlib.Path]:
    """Headers of the reference tree that `path` includes, transitively."""
    out = []
    for inc in re.findall(r'^\s*#\s*include\s*"([^"]+)"', path.read_text(errors="replace"), flags=re.M):
        h = (path.parent / inc).resolve()
        if h.suffix == ".h" and h.is_file() and h not in seen and REF.resolve() in h.parents:
            seen.add(h)
            out.append(h)
            out += local_includes(h, seen)
    return out


def main_sources() -> list[str]:
    mk = (REF / "Makefile.am").read_text()
    m = re.search(r"__common_sources\s*=(.*?)\n\s*\n", mk, flags=re.S)
    files = re.findall(r"([\w/]+\.c)", m.group(1))
    return ["OpenAcc/main.c"] + files


def generated_from() -> set[str]:
    """Files the single-precision generator converts (names suffixed _f)."""
    src = (REF / "double_to_single_transformer.py").read_text()
    m = re.search(r"fileNames\s*=\s*\[(.*?)\]", src, flags=re.S)
    return {f for f in re.findall(r"'([\w/]+\.c)'", m.group(1))}


def renamed_names(conv: set[str]) -> set[str]:
    """Function names the generator suffixes with _f: those defined in the
    files it converts (a call to any other function keeps its name)."""
    names: set[str] = set()
    for f in conv:
        names |= set(definitions(REF / f))
    return names


def main() -> int:
    if not REF.is_dir():
        print("reference tree not found", file=sys.stderr)
        return 1
    sources = main_sources()
    kept = [f for f in sources if f not in DROPPED_FILES]
    conv = generated_from()
    renamed = renamed_names(conv)
    # functions removed from the program, with the file that defined them
    removed: dict[str, str] = {}
    for f in DROPPED_FILES:
        for name, st in definitions(REF / f, variables=True).items():
            if not st:
                removed.setdefault(name, f)
    for f, names in DROPPED_DEFS.items():
        d = definitions(REF / f)
        for name in names:
            assert name in d and not d[name], (f, name)
            removed[name] = f
    own = {f: definitions(REF / f, variables=True) for f in kept}
    needed: dict[str, list[str]] = {}
    for f in kept:
        # the file's own bodies plus the static inline bodies of the headers it includes
        calls = body_words(strip_c((REF / f).read_text(errors="replace")))
        for h in local_includes(REF / f, set()):
            for k, v in body_words(strip_c(h.read_text(errors="replace"))).items():
                calls[k] = calls.get(k, 0) + v
        for name in sorted(removed):
            if name in own[f] and name not in DROPPED_DEFS.get(f, []):
                continue  # the kept file has its own definition
            if calls.get(name, 0) > 0:
                needed.setdefault(name, []).append(f)
                if f in conv and f not in DROPPED_TWINS_OF_KEPT and name in renamed:
                    # its generated twin calls name_f
                    needed.setdefault(name + "_f", []).append("sp_" + f.split("/")[-1] + " (generated)")
            if calls.get(name + "_f", 0) > 0:
                needed.setdefault(name + "_f", []).append(f)
    out = {
        "about": "functions defined in the reference files the library replaces, that the kept reference "
                 "files call (name -> callers); generated by tests/golden/make_abi_closure.py",
        "dropped_files": DROPPED_FILES,
        "dropped_definitions": DROPPED_DEFS,
        "dropped_twins_of_kept_files": DROPPED_TWINS_OF_KEPT,
        "needed": {k: sorted(set(v)) for k, v in sorted(needed.items())},
        "removed_definitions": dict(sorted(removed.items())),
    }
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(removed)} removed definitions, {len(needed)} still called -> {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
