"""The reference's interchange formats.

- Hamiltonian text: SumOps' operator<< (inc/ops.h:588-623, inc/cpuimpl.h:
  478-506) -- terms joined by " + " (or " - " before a term whose text starts
  with a minus), each term "coef*sx(0).sz(3)" with the coefficient in the
  standard sigma_y convention written as "i", "-i", "(re+imi)", "re" or
  "imi" (omitted for 1, "-" for -1), "Id" for the identity and "0" for an
  empty sum; the SYK driver writes "H_LL = ..." and "H_RR = ..." lines
  (examples/syk.cc:429-438).  ham_to_string produces it through the C++ API
  and parse_ham_string reads it back (numbers carry the ostream's 6
  significant digits).
- Result files: one "t Re Im" line per time point (examples/syk.cc:230-235).
- The JSON service: reapers_service reads one request from standard input and
  writes the reply (tests/service.cc:21-240); service_request runs it.
"""
import ctypes as _c
import json
import os
import re
import subprocess

import numpy as np

from . import HERE, host, hcheck, _ptr, _i64

SERVICE_EXE = os.path.join(HERE, "reapers_service")

_SITE_CODE = {"sx": 1, "sz": 2, "sy": 3}


def _text_call(fn):
    cap = 1 << 16
    while True:
        buf = _c.create_string_buffer(cap)
        n = _i64()
        fn(buf, cap, _c.byref(n))
        if n.value < cap:
            return buf.value.decode()
        cap = n.value + 1


def ham_to_string(bits, coeff, fp=64):
    """SumOps' operator<< of a term list (bits, complex coefficients in the
    internal convention, where sigma_y carries its i)."""
    bits = np.ascontiguousarray(bits, np.uint64)
    co = np.empty(2 * len(bits))
    co[0::2] = np.real(coeff)
    co[1::2] = np.imag(coeff)
    return _text_call(lambda b, cap, n: hcheck(host().rph_ham_to_string(
        fp, _ptr(bits), _ptr(co), len(bits), b, cap, n)))


def syk_block_dump(N, sp=0.0, J=1.0, seed=0, regularize=False, fp=32):
    """The SYK driver's Hamiltonian dump ("H_LL = ..." and "H_RR = ..."
    lines) of SYKBlockForm(N, sp, J).gen_ham(std::mt19937(seed))."""
    return _text_call(lambda b, cap, n: hcheck(host().rph_syk_block_dump(
        fp, N, sp, J, seed, int(regularize), b, cap, n)))


def _parse_coef(text):
    if text == "i":
        return 1j
    if text == "-i":
        return -1j
    m = re.fullmatch(r"\((.+)\+(.+)i\)", text)
    if m:
        return complex(float(m.group(1)), float(m.group(2)))
    if text.endswith("i"):
        return complex(0.0, float(text[:-1]))
    return complex(float(text), 0.0)


def _parse_term(text):
    if text == "0":
        return None
    if "*" in text:
        ctext, otext = text.split("*", 1)
        coef = _parse_coef(ctext)
    elif text.startswith("-"):
        coef, otext = -1.0 + 0j, text[1:]
    else:
        coef, otext = 1.0 + 0j, text
    bits, nsy = 0, 0
    if otext != "Id":
        for part in otext.split("."):
            m = re.fullmatch(r"(s[xyz])\((\d+)\)", part)
            if not m:
                raise ValueError("bad site operator %r" % part)
            code = _SITE_CODE[m.group(1)]
            bits |= code << (2 * int(m.group(2)))
            nsy += code == 3
    # the text holds the standard-convention coefficient; the internal one
    # carries i for each sigma_y (inc/ops.h:366-377)
    return bits, coef * (1j ** (nsy % 4))


def parse_ham_string(text):
    """Term list (bits, internal-convention coefficients) of a SumOps text."""
    text = text.strip()
    if text.startswith("H_LL = ") or text.startswith("H_RR = "):
        text = text[7:]
    pieces = re.split(r" ([+-]) ", text)
    terms = [pieces[0]]
    for sign, body in zip(pieces[1::2], pieces[2::2]):
        terms.append(body if sign == "+" else "-" + body)
    bits, coeff = [], []
    for t in terms:
        r = _parse_term(t)
        if r is not None:
            bits.append(r[0])
            coeff.append(r[1])
    return np.array(bits, np.uint64), np.array(coeff, np.complex128)


def format_result_lines(t, v, digits=6):
    """Lines 't Re Im', one per time point, as the reference's result files
    hold them (an ostream's default format: %g with 6 significant digits;
    digits=17 keeps every bit of a double)."""
    fmt = "%%.%dg" % digits
    return ["%s %s %s" % (fmt % float(tk), fmt % float(np.real(vk)), fmt % float(np.imag(vk)))
            for tk, vk in zip(t, v)]


def format_results_cpp(dt, v):
    """The C++ driver's write_results of v at t = k dt (for comparison)."""
    vv = np.empty(2 * len(v))
    vv[0::2] = np.real(v)
    vv[1::2] = np.imag(v)
    return _text_call(lambda b, cap, n: hcheck(host().rph_format_results(
        dt, _ptr(vv), len(v), b, cap, n))).splitlines()


def write_result_file(path, t, v, digits=6):
    with open(path, "w") as f:
        for line in format_result_lines(t, v, digits):
            f.write(line + "\n")


def read_result_file(path):
    """(t, v) from a 't Re Im' result file."""
    rows = np.loadtxt(path, ndmin=2)
    return rows[:, 0], rows[:, 1] + 1j * rows[:, 2]


def service_request(request, params, timeout=300):
    """Run reapers_service on {"request": request, "params": params} and
    return the parsed reply; raises RuntimeError with its standard error on
    failure."""
    if not os.path.exists(SERVICE_EXE):
        raise RuntimeError("reapers_service is not built: run __graft_entry__.build()")
    req = json.dumps({"request": request, "params": params})
    r = subprocess.run([SERVICE_EXE], input=req, capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("reapers_service failed (%d): %s" % (r.returncode, r.stderr.strip()))
    return json.loads(r.stdout)


def cmat(rows_pairs):
    """A row-major list of [re, im] pairs as a square complex matrix."""
    a = np.asarray(rows_pairs, np.float64)
    n = int(round(np.sqrt(len(a))))
    assert n * n == len(a)
    return (a[:, 0] + 1j * a[:, 1]).reshape(n, n)


def cvec(pairs):
    a = np.asarray(pairs, np.float64)
    return a[:, 0] + 1j * a[:, 1]
