"""ctypes mirrors of the C-ABI structs declared in include/*.h."""
import ctypes

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)


class LsmMechanism(ctypes.Structure):
    """include/lsm_mechanism.h: struct lsm_mechanism."""
    _fields_ = [
        ("ns", ctypes.c_int32), ("nr", ctypes.c_int32),
        ("nasa", _d), ("lnA", _d), ("beta", _d), ("E_over_R", _d), ("rflags", _i),
        ("reac_ptr", _i), ("reac_sp", _i), ("reac_nu", _d),
        ("prod_ptr", _i), ("prod_sp", _i), ("prod_nu", _d),
        ("ntb", ctypes.c_int32), ("tb_rxn", _i), ("tb_ptr", _i), ("tb_sp", _i), ("tb_eff_m1", _d),
        ("nfo", ctypes.c_int32), ("fo_rxn", _i), ("fo_type", _i), ("fo_inf", _d), ("fo_troe", _d),
        ("fo_single", _i), ("fo_ptr", _i), ("fo_sp", _i), ("fo_eff_m1", _d), ("fo_sri", _d),
        ("ord_ptr", _i), ("ord_sp", _i), ("ord_val", _d), ("rev_arr", _d),
        ("npl", ctypes.c_int32), ("pl_rxn", _i), ("pl_ptr", _i), ("pl_par", _d),
        ("nch", ctypes.c_int32), ("ch_rxn", _i), ("ch_dims", _i), ("ch_lim", _d), ("ch_ptr", _i), ("ch_a", _d),
    ]


class LsmTransport(ctypes.Structure):
    """include/lsm_transport.h: struct lsm_transport."""
    _fields_ = [
        ("ns", ctypes.c_int32), ("mw", _d), ("lambda_fit", _d), ("eta_fit", _d), ("gamma_fit", _d),
        ("nteta", ctypes.c_int32), ("teta_sp", _i), ("teta_fit", _d), ("viscosity_model", ctypes.c_int32),
    ]


class LsmTransportOptions(ctypes.Structure):
    """include/lsm_transport.h: struct lsm_transport_options."""
    _fields_ = [
        ("diffusivity_model", ctypes.c_int32), ("lewis", _d), ("simplified", ctypes.c_int32),
        ("soret", ctypes.c_int32), ("soot_reference", ctypes.c_int32), ("n_soot", ctypes.c_int32),
        ("soot_sp", _i), ("soot_corr", _d),
    ]


class LsmOdeParameters(ctypes.Structure):
    """include/laminarsmoke_amd.h: struct lsm_ode_parameters."""
    _fields_ = [
        ("type", ctypes.c_int32),
        ("relative_tolerance", ctypes.c_double),
        ("absolute_tolerance", ctypes.c_double),
        ("minimum_step", ctypes.c_double),
        ("maximum_step", ctypes.c_double),
        ("initial_step", ctypes.c_double),
        ("maximum_order", ctypes.c_int32),
        ("maximum_number_of_steps", ctypes.c_int32),
        ("full_pivoting", ctypes.c_int32),
        ("linear_algebra", ctypes.c_char * 32),
    ]


class LsmReactorOptions(ctypes.Structure):
    """include/laminarsmoke_amd.h: struct lsm_reactor_options."""
    _fields_ = [
        ("pressure", ctypes.c_double),
        ("min_temperature", ctypes.c_double),
        ("energy_equation", ctypes.c_int32),
        ("mass_fractions_tol", ctypes.c_double),
        ("constant_volume", ctypes.c_int32),
        ("strang_compact", ctypes.c_int32),
    ]


N_COUNTERS = 12
COUNTER_NAMES = (
    "steps", "function_calls", "function_calls_jacobian", "jacobians", "factorizations",
    "linear_solutions", "convergence_failures", "error_test_failures", "max_order_used",
    "rhs_total", "clamp_events", "newton_iterations",
)
