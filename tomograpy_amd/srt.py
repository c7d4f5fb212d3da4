"""Reconstruction parameters kept in the output map's header
(tomograpy/srt.py:230-290), so that a map records how it was computed.

Keys are cut to the 8 characters of a FITS keyword; a parameter with several
values is stored as KEY1, KEY2, ... (first 7 characters + index); values are
stored as strings, as the reference does.
"""
from .headers import dict_to_array


def persistency_header(object_header, data_params, mask_params, opt_params):
    """srt.py:230-256: the object header plus every parameter as a string."""
    full = dict(data_params, **mask_params)
    full.update(opt_params)
    if "model" in full and callable(full["model"]):
        full["model"] = full["model"].__name__
    out = {}
    for k, v in full.items():
        if hasattr(v, "__iter__") and not isinstance(v, str):
            v = list(v)
            if len(v) == 1:
                out[k[:8].upper()] = str(v[0])
            else:
                for i, p in enumerate(v):
                    out[(k[:7] + str(i + 1)).upper()] = str(p)
        else:
            out[k[:8].upper()] = str(v)
    header = dict(object_header)
    header.update(out)
    return header


def params_from_header(h):
    """srt.py:258-290: (obj_params, data_params, opt_params, mask_params) back
    from a persistency header (values as stored, i.e. strings)."""
    obj_params = {k: dict_to_array(h, k.upper()) for k in ("naxis", "crpix", "crval", "cdelt")}
    data_params = {}
    for k in ("obj_rmin", "obj_rmax", "data_rmin", "data_rmax", "negative"):
        if k.upper()[:8] in h:
            data_params[k] = h[k.upper()[:8]]
    opt_params = {k: h[k.upper()[:8]] for k in ("model", "optimizer", "maxiter", "tol")}
    opt_params["hyperparameters"] = dict_to_array(h, "HYPERS")
    mask_params = {}
    for k in ("tmin", "tmax", "time_step", "bin_factor"):
        if k.upper()[:8] in h:
            mask_params[k] = h[k.upper()[:8]]
    for k in ("telescop", "instrument"):
        try:
            mask_params[k] = dict_to_array(h, k.upper()[:8])
        except (KeyError, ValueError):
            pass
    return obj_params, data_params, opt_params, mask_params
