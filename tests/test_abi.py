"""C-ABI checks that need no GPU: the library loads, exports every declared entry
point, and its host-only functions (defaults, dictionary parsing) behave like
OpenSMOKE::ODE_Parameters and laminarSMOKE's readOptions.H."""
import ctypes
import os
import re

import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd._abi import LsmOdeParameters, LsmReactorOptions, LsmMechanism, LsmTransport, LsmTransportOptions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lsm_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = api.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert lib.lsm_version().startswith(b"laminarsmoke_amd")


def test_struct_layouts_match_header():
    """ctypes mirrors have the sizes the C compiler gives the header structs."""
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "laminarsmoke_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lsm_mechanism), sizeof(lsm_ode_parameters),
         sizeof(lsm_reactor_options), offsetof(lsm_mechanism, fo_eff_m1), offsetof(lsm_ode_parameters, linear_algebra),
         offsetof(lsm_mechanism, rev_arr), offsetof(lsm_mechanism, npl), offsetof(lsm_mechanism, ch_a),
         sizeof(lsm_transport), offsetof(lsm_transport, viscosity_model), sizeof(lsm_transport_options),
         offsetof(lsm_transport_options, soot_corr));
  return 0; }
'''
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).split()
    got = [int(x) for x in out]
    assert got == [ctypes.sizeof(LsmMechanism), ctypes.sizeof(LsmOdeParameters), ctypes.sizeof(LsmReactorOptions),
                   LsmMechanism.fo_eff_m1.offset, LsmOdeParameters.linear_algebra.offset,
                   LsmMechanism.rev_arr.offset, LsmMechanism.npl.offset, LsmMechanism.ch_a.offset,
                   ctypes.sizeof(LsmTransport), LsmTransport.viscosity_model.offset,
                   ctypes.sizeof(LsmTransportOptions), LsmTransportOptions.soot_corr.offset]


def test_defaults_mirror_ode_parameters():
    """ODE_Parameters::ODE_Parameters() (ODE_Parameters.hpp:239-256)."""
    p = api.ODE_Parameters()
    assert p.relative_tolerance() == pytest.approx(100 * 1.1920928955078125e-07)
    assert p.absolute_tolerance() == 1e-10
    assert not p.full_pivoting()
    assert p.maximum_order() == -1
    assert p.type() == api.ODE_INTEGRATOR["OpenSMOKE_GPU"]
    assert p.c.linear_algebra == b"Eigen"


def test_parse_openfoam_subdictionary():
    """LS/readOptions.H:474-530 keys, as written in the tutorials' constant/solverOptions."""
    p = api.ODE_Parameters()
    p.SetupFromDictionary("""
        OdeHomogeneous
        {
            odeSolver       "OpenSMOKE_GPU";   // the GPU integrator
            relTolerance    1e-6;
            absTolerance    1e-12;
            maximumOrder    5;
            fullPivoting    true;
            /* CHEMEQ2 sub-dictionary is not read for this solver */
        }""")
    assert p.type() == 100
    assert p.relative_tolerance() == 1e-6 and p.absolute_tolerance() == 1e-12
    assert p.full_pivoting() and p.maximum_order() == 5
    p.SetupFromDictionary("OdeHomogeneous { odeSolver OpenSMOKE; fullPivoting off; }")
    assert p.type() == 0 and not p.full_pivoting()


def test_parse_opensmoke_keywords():
    """@-keywords of ODE_Parameters::SetupFromDictionary (grammar ODE_Parameters.h:163-239)."""
    p = api.ODE_Parameters()
    p.SetupFromDictionary("@OdeSolver OpenSMOKE_GPU; @RelativeTolerance 1e-7; @AbsoluteTolerance 1e-13;"
                          "@MaximumNumberOfSteps 1000; @FullPivoting true; @LinearAlgebra Eigen;"
                          "@SparseSolver none; @MaximumOrder 3;")
    assert p.relative_tolerance() == 1e-7 and p.absolute_tolerance() == 1e-13
    assert p.c.maximum_number_of_steps == 1000 and p.full_pivoting() and p.maximum_order() == 3


@pytest.mark.parametrize("bad", [
    "odeSolver Foo;",                      # unknown solver (readOptions.H:495-504)
    "relTolerance abc;",                   # malformed number
    "fullPivoting maybe;",                 # not an OpenFOAM Switch word
    "unknownKey 1;",
    "OdeHomogeneous { relTolerance 1e-6;",  # unbalanced
    "relTolerance;",
])
def test_parse_errors_are_reported_not_fatal(bad):
    p = api.ODE_Parameters()
    before = p.relative_tolerance()
    with pytest.raises(api.LsmError):
        p.SetupFromDictionary(bad)
    assert p.relative_tolerance() == before           # unchanged on error
    assert api.lib().lsm_last_error() != b""


def test_null_arguments_rejected():
    L = api.lib()
    assert L.lsm_ode_parameters_default(None) == -1
    assert L.lsm_integrate_batch(None, 1, 0.0, 101325.0, None, None, None, None, None, None) == -1
    assert L.lsm_destroy(None) == 0
    out = ctypes.c_void_p()
    assert L.lsm_create(None, None, None, ctypes.byref(out)) == -1
