"""libhh loads and exports every symbol that include/*.h declare; the
pybind11 module exposes the reference's surface (no device calls here)."""
import ctypes

from hignn_amd import abi


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [s for s in abi.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert len(abi.header_symbols()) >= 25


def test_pybind_surface_matches_reference():
    from hignn_amd import hignn
    # interface/HignnInterface.cpp:14-64
    assert callable(hignn.Init) and callable(hignn.Finalize)
    for meth in ["load_two_body_model", "load_three_body_model", "set_epsilon", "set_eta",
                 "set_max_iter", "set_mat_pool_size_factor", "set_max_far_dot_work_node_size",
                 "set_max_relative_coord", "set_post_check_flag", "set_use_symmetry_flag",
                 "set_max_far_field_distance", "update_coord", "dot", "dense_dot"]:
        assert hasattr(hignn.HignnModel, meth), meth
    for cls in (hignn.ExplicitEuler, hignn.ExplicitRk4):
        for meth in ["set_time_step", "set_final_time", "set_num_rigid_body",
                     "set_velocity_func", "set_output_func", "set_x_lim", "set_y_lim",
                     "set_z_lim", "set_output_step", "initialize", "run"]:
            assert hasattr(cls, meth), (cls, meth)
    assert hasattr(hignn.ExplicitRk4, "set_threshold")


def test_errors_are_reported():
    lib = abi.lib()
    h = ctypes.c_void_p()
    rc = lib.hh_tree_build(None, 10, 64, 1, 500000, ctypes.byref(h))
    assert rc == -1
    assert b"NULL" in lib.hh_last_error()


def test_null_handles_and_bad_sizes_are_rejected_before_any_device_call():
    # Every family of entry points checks its handle and arguments first and
    # reports HH_ERR_ARG (-1) or HH_ERR_STATE (-2) with a message.
    lib = abi.lib()
    P, I64, F, D, I = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_double,
                       ctypes.c_int)
    buf = (ctypes.c_float * 16)()
    b3 = ctypes.cast(buf, P)
    rows = ctypes.c_int()
    cases = [
        ("hh_model_dot_device", (P, P, P, I64, P), (None, b3, b3, 4, None)),
        ("hh_model_dot_enqueue", (P, P, P, I64, P), (None, b3, b3, 4, None)),
        ("hh_model_tune_far_field", (P, P, I64, P, P, I, P),
         (None, b3, 4, None, None, 0, ctypes.cast(ctypes.byref(rows), P))),
        ("hh_forcing_set_body", (P, P), (None, b3)),
        ("hh_stepper_set", (P, D, D, D), (None, 0.0, 1.0, 0.0)),  # non-positive step
        ("hh_stepper_set", (P, D, D, D), (None, 0.1, 1.0, 0.0)),
        ("hh_stepper_run", (P, I64), (None, 0)),
        ("hh_three_body_velocity_device", (P, P, P, I64, P, I, P), (None, b3, b3, 4, b3, 0, None)),
        ("hh_surrogate_eval_device", (P, P, I64, P), (None, b3, 4, b3)),
        ("hh_surrogate_create", (F, P), (1.0, None)),
        ("hh_surrogate_create", (F, P), (-1.0, b3)),  # r0 must be positive
    ]
    for name, argtypes, args in cases:
        fn = ctypes.CFUNCTYPE(ctypes.c_int, *argtypes)((name, lib))  # private prototype
        rc = fn(*args)
        assert rc in (-1, -2), (name, args, rc)
        assert lib.hh_last_error(), name
