"""Masks the run-dependent values of a `picomax` stdout (test helper).

The reference's main prints the date, host, core count, memory, working
directory and wall times (src/pmx.cpp:84-203); those differ between runs and
machines, so they are replaced by placeholders before two outputs are
compared line by line.  Everything else, including the order and text of the
progress lines, must match.
"""
import re

_RULES = [
    (re.compile(r"^(Mon|Tue|Wed|Thu|Fri|Sat|Sun) \w{3} [ \d]\d \d\d:\d\d:\d\d \d{4}$"), "<date>"),
    (re.compile(r"^(  Running on ).*$"), r"\1<host>"),
    (re.compile(r"^(  Using )\d+( cores)$"), r"\1<n>\2"),
    (re.compile(r"^(  (Total|Avail) memory: )\d+( MB)$"), r"\1<n>\3"),
    (re.compile(r"^(  Working dir: ).*$"), r"\1<dir>"),
    (re.compile(r"^(working directory: ).*$"), r"\1<dir>"),
    (re.compile(r"^(\s*Elapsed time: )\S+( s)$"), r"\1<t>\2"),
]


def mask(text, golden_dir=None):
    if golden_dir:
        text = text.replace(golden_dir, "{golden}")
    out = []
    for line in text.splitlines():
        for rx, rep in _RULES:
            line = rx.sub(rep, line)
        out.append(line)
    return out


def solver_part(lines):
    """The lines from the solver heading on (the part chi_LL / chi_xyz runs share
    with the current sources; the older snapshot's banner differs)."""
    for i, l in enumerate(lines):
        if l.startswith("Susceptibility tensor matrix solver") or l.startswith("Electronic bandstructure solver"):
            return lines[i:]
    return []
