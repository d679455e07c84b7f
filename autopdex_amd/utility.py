"""DOF-structure helpers (host, numpy) with the semantics of autopdex.utility.

dict_flatten concatenates dict fields in key order (utility.py:104-128);
reshape_as is its inverse (:130-199); mask_op implements modes 'set', 'add',
'apply' on arrays or dicts (:283-368); dof_select builds per-DOF masks from a
node mask (:448-462).
"""
import numpy as np


def dict_zeros_like(arr, dtype=np.float64):
    if isinstance(arr, dict):
        return {k: np.zeros(np.shape(v), dtype=dtype) for k, v in arr.items()}
    return np.zeros(np.shape(arr), dtype=dtype)


def dict_ones_like(arr, dtype=np.float64):
    if isinstance(arr, dict):
        return {k: np.ones(np.shape(v), dtype=dtype) for k, v in arr.items()}
    return np.ones(np.shape(arr), dtype=dtype)


def dict_flatten(x):
    if isinstance(x, dict):
        return np.concatenate([np.asarray(v).ravel() for v in x.values()])
    return np.asarray(x).ravel()


def reshape_as(flat, like):
    flat = np.asarray(flat)
    if isinstance(like, dict):
        out, o = {}, 0
        for k, v in like.items():
            n = int(np.size(v))
            out[k] = flat[o:o + n].reshape(np.shape(v))
            o += n
        return out
    return flat.reshape(np.shape(like))


def mask_op(array, selection, values=None, mode="set", ufunc=None):
    if isinstance(array, dict):
        return {k: mask_op(array[k], selection[k], None if values is None else values[k], mode, ufunc)
                for k in array}
    a = np.asarray(array).ravel().copy()
    s = np.asarray(selection, dtype=bool).ravel()
    if mode == "apply":
        a = np.where(s, ufunc(a), a)
    elif mode == "set":
        a = np.where(s, dict_flatten(values), a)
    elif mode == "add":
        a = np.where(s, a + np.asarray(values).ravel(), a)
    else:
        raise ValueError(f"Invalid mode '{mode}'.")
    return a.reshape(np.shape(array))


def mask_select(array, selection):
    if isinstance(array, dict):
        return {k: mask_select(array[k], selection[k]) for k in array}
    return np.where(np.asarray(selection, dtype=bool).ravel(), np.asarray(array).ravel(), 0).reshape(np.shape(array))


def dof_select(dirichlet_nodes, selected_fields):
    if isinstance(selected_fields, bool):
        return np.asarray(dirichlet_nodes) * selected_fields
    return np.outer(dirichlet_nodes, np.asarray(selected_fields))
