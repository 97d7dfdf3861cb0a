# Relinks the reference's own drivers (compiled from their sources where they
# lie) against this library's mpcd:: API instead of the reference's library:
# the drop-in check.  Outputs go to oracle/_ref/dropin/ (git-ignored; they
# travel with the tree to the GPU host, where the GPU tests run them).
#
#   make -C oracle -f dropin.mk
REF ?= /root/reference
ROOT := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT := $(ROOT)/oracle/_ref/dropin
LIBDIR := $(ROOT)/src_incompressible_openmp_amd
LIB := $(LIBDIR)/libmpcd_gpu.so
CXX ?= g++
FLAGS := -std=c++17 -O2 -fopenmp -I$(ROOT)/include/mpcd -I$(ROOT)/include

DRIVERS := $(notdir $(basename $(wildcard $(REF)/src/main_*.cpp)))

all: $(addprefix $(OUT)/,$(DRIVERS))

$(OUT)/%: $(REF)/src/%.cpp $(LIB) $(wildcard $(ROOT)/include/mpcd/*.h)
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) $< -L$(LIBDIR) -lmpcd_gpu -Wl,-rpath,'$$ORIGIN/../../../src_incompressible_openmp_amd' -o $@

.PHONY: all
