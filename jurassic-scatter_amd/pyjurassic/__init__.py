"""Python side of the MI355X forward model: ctypes bindings of the C ABI
(include/jurassic_amd.h), struct mirrors and the synthetic-input generator."""
