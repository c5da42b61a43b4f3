# Builds the reference's own CPU-only code into oracle/_ref/ (test
# infrastructure).  Sources are compiled where they lie; nothing is copied.
# Only the CPU headers compile without CUDA (lmo.cuh and the solver need
# CUDA headers, so the GPU path is not built here).
REF      ?= /root/reference
CXX      ?= g++
CXXFLAGS := -O2 -std=c++17 -fPIC -pthread -I$(REF)/include
OUT      := _ref

all: $(OUT)/test_cpu $(OUT)/libref.so

$(OUT):
	mkdir -p $(OUT)

# The reference's CPU known-answer tests (tests/test_cpu.cpp).
$(OUT)/test_cpu: $(REF)/tests/test_cpu.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) $< -o $@

# C-ABI shim over the reference's CPU helpers.
$(OUT)/libref.so: ref_shim.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -shared ref_shim.cpp -o $@

.PHONY: all
