"""tomograpy_amd: TomograPy's Siddon conic projector/backprojector on MI355X.

The names below are the reference's top-level surface (tomograpy/__init__.py
re-exports siddon.* and, when ``lo`` is importable, lo_wrapper.*).
"""
from .infoarray import InfoArray, asinfoarray
from .headers import (
    INF, bitpix, bitpix_inv, rotation_matrix, header_rotation_matrix, full_rotation_matrix,
    map_borders, centered_cubic_map_header, centered_cubic_map, centered_image_header,
    centered_image, centered_stack, fov, circular_trajectory_data, circular_trajectory_header,
    dataarray_from_header, fitsarray_from_header, object_from_header, array_to_dict,
    dict_to_array, update_observer,
)
from .siddon import (
    projector, backprojector, projector4d, backprojector4d, conic_image_projector,
    conic_image_backprojector, ray_projector, check_projector_inputs, C_full_unit_vector,
    C_full_intersection_parameters, count_steps, check_status,
)
from . import solar, srt
from .solar import read_data, update_header

version = "0.1.0"
