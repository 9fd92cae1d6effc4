"""The programmatic scene builder lives in the package (mypbr_amd.scenebuild)."""
from mypbr_amd.scenebuild import DescBuilder, camera, roughness_to_alpha  # noqa: F401
