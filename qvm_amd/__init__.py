"""qvm_amd -- MI355X (gfx950) statevector engine for quil-lang/qvm's pure-state path.

The engine is a C-ABI HIP library (include/qvmhip.h, sources in qvm_amd/csrc,
built into qvm_amd/_lib/libqvmhip.so).  This package mirrors qvm's Lisp API
on top of it: ``make_qvm``, ``load_program``, ``run``, ``measure``,
``measure_all``, ``amplitudes`` (qvm_amd.qvm) and the state protocol
(qvm_amd.state.HipPureState).
"""
from .gates import gate_matrix, nat_tuple  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # Lazy: importing the package does not load the HIP library.
    if name in ("make_qvm", "run_program", "PureStateQVM", "apply_gate_to_state"):
        from . import qvm
        return getattr(qvm, name)
    if name in ("HipPureState", "Program"):
        from . import state
        return getattr(state, name)
    raise AttributeError(name)
