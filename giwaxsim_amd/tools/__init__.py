"""Reference-compatible call surface (tools/voxelgrids.py, tools/detector.py,
tools/comparison.py, tools/utilities.py, tools/ptable_dict.py of GIWAXSim)."""
