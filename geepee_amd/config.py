"""Constants of the AEP path (geepee/config.py:11-16), and the launch switches of
this package."""
import os

JITTER = 1e-5
GH_DEGREE = 10
PROP_MM = 'MM'
PROP_LIN = 'LIN'
PROP_MC = 'MC'
MC_NO_SAMPLES = 5

# SDGPR forks each layer's independent chains (q(u) set-up, hyper-parameter adjoint)
# onto a stream of its own; GEEPEE_AMD_STREAMS=0 keeps every launch on one stream.
MULTI_STREAM = os.environ.get('GEEPEE_AMD_STREAMS', '1') != '0'
