"""Traced / learned-stress stage path (jathenax/hydro.py:111-119, 134-142).

Not built yet: the fused stage kernel keeps face states and fluxes on-chip, so the
traced dictionaries and the per-stage predictor call need the unfused debug kernels.
"""


def traced_step(*args, **kwargs):
    raise NotImplementedError("trace=True and predict_stress need the unfused debug stage "
                              "path, which is not built yet")


def evolve_with_predictor(*args, **kwargs):
    raise NotImplementedError("predict_stress needs the unfused debug stage path, which is "
                              "not built yet")
