"""The reference-side Lisp binding (lisp/qvm-hip.lisp) read as text: no Lisp
exists in this image, so these checks stand in for loading it.

* every cffi:defcfun names an entry point of include/qvmhip.h that the built
  library exports, with the header's number of arguments;
* apply-gate-to-state has a method on hip-pure-state for every gate class
  the reference dispatches on for pure states (src/apply-gate.lisp:109-181);
* the state protocol (src/state-representation.lisp:25-47) and measurement
  protocol (src/measurement.lisp:7-143) generics have methods on it;
* the two machine methods that touch the host vector (reset-quantum-state,
  src/qvm.lisp:84-87; the compiled projector's transition,
  src/transition.lisp:188-194) are routed for it;
* the machine is the reference's pure-state-qvm, made with make-instance
  (make-qvm has no :state key, src/qvm.lisp:150-164);
* the forms are balanced.
"""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LISP = os.path.join(REPO, "lisp", "qvm-hip.lisp")
HEADER = os.path.join(REPO, "include", "qvmhip.h")

# Gate classes with a pure-state apply-gate-to-state method in the reference
# (src/apply-gate.lisp:109-181).
GATE_CLASSES = ["quil:simple-gate", "quil:parameterized-gate", "quil:sequence-gate",
                "quil:permutation-gate", "quil:controlled-gate", "quil:forked-gate",
                "quil:dagger-gate", "qvm::compiled-matrix-gate-application",
                "qvm::compiled-inlined-matrix-gate-application",
                "qvm::compiled-permutation-gate-application"]

PROTOCOL = [r"qvm::num-qubits \(\(state hip-pure-state\)\)",
            r"qvm:state-elements \(\(state hip-pure-state\)\)",
            r"\(setf qvm:state-elements\) \(new-value \(state hip-pure-state\)\)",
            r"qvm:set-to-zero-state \(\(state hip-pure-state\)\)",
            r"qvm::requires-swapping-amps-p \(\(state hip-pure-state\)\)",
            r"qvm::swap-internal-amplitude-pointers \(\(state hip-pure-state\)\)",
            r"qvm::check-allocate-computation-space \(\(state hip-pure-state\)\)",
            r"qvm::get-excited-state-probability \(\(state hip-pure-state\) qubit\)",
            r"qvm:force-measurement \(measured-value qubit \(state hip-pure-state\) excited-probability\)",
            r"qvm::measure-all-state \(\(state hip-pure-state\) \(qvm qvm::base-qvm\)\)",
            r"qvm::reset-quantum-state :around \(\(qvm qvm::base-qvm\)\)",
            r"qvm:transition :around \(\(qvm qvm:pure-state-qvm\) \(instr qvm::compiled-measurement\)\)"]


def lisp_text():
    with open(LISP) as f:
        return f.read()


def strip_comments_and_strings(src: str) -> str:
    out, i, n = [], 0, len(src)
    while i < n:
        c = src[i]
        if c == ";":
            while i < n and src[i] != "\n":
                i += 1
        elif c == '"':
            i += 1
            while i < n and src[i] != '"':
                i += 2 if src[i] == "\\" else 1
            i += 1
            out.append('""')
        elif c == "#" and i + 1 < n and src[i + 1] == "\\":
            i += 3  # a character literal such as #\(
            out.append("x")
        else:
            out.append(c)
            i += 1
    return "".join(out)


def top_level_forms(code: str):
    forms, depth, start = [], 0, None
    for i, c in enumerate(code):
        if c == "(":
            if depth == 0:
                start = i
            depth += 1
        elif c == ")":
            depth -= 1
            assert depth >= 0, f"unbalanced ')' at offset {i}"
            if depth == 0:
                forms.append(code[start:i + 1])
    assert depth == 0, "unbalanced '(' at end of file"
    return forms


def header_params():
    with open(HEADER) as f:
        h = f.read()
    out = {}
    for m in re.finditer(r"QVMHIP_API\s+[\w\s\*]+?\b(qvmhip_\w+)\s*\(([^;]*?)\)\s*;", h, re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    return out


def defcfuns():
    code = strip_comments_and_strings(lisp_text())
    src = lisp_text()
    out = {}
    for m in re.finditer(r'\(cffi:defcfun \("(qvmhip_\w+)" [%\w-]+\) :[\w-]+', src):
        # the argument specs follow the return type up to the form's end
        depth, i = 1, m.start() + 1
        while depth:
            c = src[i]
            depth += (c == "(") - (c == ")")
            i += 1
        form = src[m.start():i]
        body = form[form.index(")", form.index('"')) + 1:]
        out[m.group(1)] = len(re.findall(r"\(\w[\w-]* :[\w-]+\)", body))
    assert code  # the stripped text parses (checked in test_forms_balanced)
    return out


def test_forms_balanced():
    forms = top_level_forms(strip_comments_and_strings(lisp_text()))
    heads = [re.match(r"\(([^\s()]+)", f).group(1) for f in forms]
    assert heads[0] == "defpackage" and heads[1] == "in-package"
    assert heads.count("defmethod") >= 15


def test_defcfuns_match_header_and_library():
    from qvm_amd import _native as N
    hdr = header_params()
    cf = defcfuns()
    assert len(cf) >= 14
    lib = N.lib()
    for name, nargs in cf.items():
        assert name in hdr, f"{name} is not declared in include/qvmhip.h"
        assert hdr[name] == nargs, f"{name}: {nargs} arguments in the binding, {hdr[name]} in the header"
        assert hasattr(lib, name), f"{name} is not exported by libqvmhip.so"


def test_apply_gate_methods_for_every_reference_gate_class():
    src = lisp_text()
    macro_call = src[src.index("(define-device-gate-methods\n"):]
    macro_call = macro_call[:macro_call.index("))\n") + 2]
    covered = set(re.findall(r"\((quil:[\w-]+) (?:t|nil)\)", macro_call))
    covered |= set(re.findall(r"\(defmethod qvm:apply-gate-to-state \(\(gate ([\w:-]+)\)\s+\(state hip-pure-state\)",
                              src))
    assert covered == set(GATE_CLASSES), sorted(set(GATE_CLASSES) ^ covered)
    # the macro's methods specialise the state argument on the device class
    assert "((gate ,class) (state hip-pure-state) qubits" in src


@pytest.mark.parametrize("pattern", PROTOCOL)
def test_protocol_methods_present(pattern):
    assert re.search(r"\(defmethod " + pattern, lisp_text()), pattern


def test_state_is_a_pure_state_and_machine_is_pure_state_qvm():
    src = lisp_text()
    assert re.search(r"\(defclass hip-pure-state \(qvm:pure-state\)", src)
    assert "(make-instance 'qvm:pure-state-qvm" in src
    assert ":state (make-hip-pure-state" in src
    assert "(make-qvm" not in strip_comments_and_strings(src)
    # the handle is set after make-instance (pure-state's :after method must
    # see the placeholder), and the finalizer does not close over the state
    assert re.search(r"\(setf \(slot-value state 'handle\) h\)\s+\(tg:finalize state \(lambda \(\) \(%destroy h\)\)\)",
                     src)
