"""exmarg_amd: the exmarg segment analysis (exmarg.search) on AMD MI355X GPUs through a HIP library."""

__version__ = "0.1.0"
