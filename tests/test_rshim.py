"""CPU checks of the Rcpp shim (rwig_amd/rshim/hip_backend.cpp), which R's
absence keeps from being compiled here:

- every output list has the element names, in the order, of the reference
  function it stands in for (src/sinkhorn.cpp:28-43, :98-108, :154-169;
  src/barycenter.cpp:43-59, :125-135, :190-206; src/wdl.cpp:48-49, :96-97);
- every call into librwig_hip.so names a function include/rwig_hip.h declares,
  with the declared number of arguments.
"""
import os
import re

from rwig_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "rwig_amd", "rshim", "hip_backend.cpp")

EXPECTED = {
    "sinkhorn_vanilla_hip": [
        ["P", "grad_a", "u", "v", "loss", "iter", "err", "return_status"],
        ["P", "u", "v", "loss", "iter", "err", "return_status"]],
    "barycenter_parallel_hip": [
        ["b", "grad_A", "grad_w", "loss", "U", "V", "iter", "err", "return_status"],
        ["b", "U", "V", "iter", "err", "return_status"]],
    "wdl_cpp_hip": [["A", "W", "Yhat"]],
    "sinkhorn_log_hip": [
        ["P", "grad_a", "f", "g", "loss", "iter", "err", "return_status"],
        ["P", "f", "g", "loss", "iter", "err", "return_status"]],
    "barycenter_log_hip": [
        ["b", "grad_A", "grad_w", "loss", "F", "G", "iter", "err", "return_status"],
        ["b", "F", "G", "iter", "err", "return_status"]],
}


def _functions(text):
    """name -> body of each top-level function returning Rcpp::List."""
    out = {}
    for m in re.finditer(r"^Rcpp::List (\w+)\(", text, flags=re.M):
        start = text.index("{", m.end())
        depth, k = 0, start
        while True:
            depth += {"{": 1, "}": -1}.get(text[k], 0)
            if depth == 0:
                break
            k += 1
        out[m.group(1)] = text[start:k]
    return out


def _call_args(text, name):
    """Argument counts of every call name(...) in text."""
    counts = []
    for m in re.finditer(rf"\b{name}\(", text):
        depth, k, commas = 1, m.end(), 0
        while depth:
            c = text[k]
            depth += {"(": 1, ")": -1}.get(c, 0)
            if c == "," and depth == 1:
                commas += 1
            k += 1
        empty = text[m.end():k - 1].strip() == ""
        counts.append(0 if empty else commas + 1)
    return counts


def test_output_lists_match_the_reference():
    fns = _functions(open(SHIM).read())
    assert set(fns) == set(EXPECTED)
    for name, lists in EXPECTED.items():
        found = [re.findall(r'Rcpp::Named\("(\w+)"\)', blk)
                 for blk in re.findall(r"Rcpp::List::create\((.*?)\);", fns[name], flags=re.S)]
        assert found == lists, name


def test_abi_calls_match_the_header():
    header = open(_lib.HEADER).read()
    text = open(SHIM).read()
    called = set(re.findall(r"\b(rwig_\w+)\(", text))
    declared = set(_lib.header_symbols())
    assert called <= declared, called - declared
    for name in called:
        proto = re.search(rf"^\s*(?:int|const char \*|void)\s*\*?\s*{name}\((.*?)\);",
                          header, flags=re.S | re.M).group(1)
        n_decl = 0 if proto.strip() in ("", "void") else proto.count(",") + 1
        assert set(_call_args(text, name)) == {n_decl}, name
