"""Record layouts shared by the host, the HIP kernels and the oracle.

Byte layouts are those of the OpenCL structs the reference registers
(``realflare/api/tasks/opencl.py:18-61``), so arrays produced here can be
compared field by field with the reference's buffers:

* ``Ray`` (48 B): ``float3 pos`` @0, ``float3 dir`` @16, ``float2 pos_apt`` @32,
  ``float rrel`` @40, ``float reflectance`` @44.
* ``LensElement`` (64 B): ``radius, ior, height, center`` (f32) @0..12,
  ``int coating`` @16, ``float8 coefficients`` @32.
* ``Vertex`` (32 B): ``float2 pos`` @0, ``float2 uv`` @8, ``rrel`` @16,
  ``reflectance`` @20, ``intensity`` @24.
* ``Intersection`` (48 B): ``float3 pos`` @0, ``float3 normal`` @16,
  ``float incident`` @32, ``int hit`` @36.
"""
import numpy as np

LAMBDA_MIN = 390
LAMBDA_MAX = 730
LAMBDA_MID = (LAMBDA_MIN + LAMBDA_MAX) / 2

ray_dtype = np.dtype(
    {
        'names': ['pos', 'dir', 'pos_apt', 'rrel', 'reflectance'],
        'formats': [(np.float32, 3), (np.float32, 3), (np.float32, 2), np.float32, np.float32],
        'offsets': [0, 16, 32, 40, 44],
        'itemsize': 48,
    }
)

lens_element_dtype = np.dtype(
    {
        'names': ['radius', 'ior', 'height', 'center', 'coating', 'coefficients'],
        'formats': [np.float32, np.float32, np.float32, np.float32, np.int32, (np.float32, 8)],
        'offsets': [0, 4, 8, 12, 16, 32],
        'itemsize': 64,
    }
)

vertex_dtype = np.dtype(
    {
        'names': ['pos', 'uv', 'rrel', 'reflectance', 'intensity'],
        'formats': [(np.float32, 2), (np.float32, 2), np.float32, np.float32, np.float32],
        'offsets': [0, 8, 16, 20, 24],
        'itemsize': 32,
    }
)

intersection_dtype = np.dtype(
    {
        'names': ['pos', 'normal', 'incident', 'hit'],
        'formats': [(np.float32, 3), (np.float32, 3), np.float32, np.int32],
        'offsets': [0, 16, 32, 36],
        'itemsize': 48,
    }
)
