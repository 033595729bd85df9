# Builds the reference's own library probe (smoke/library_boundary_probe.cpp,
# read where it lies under the reference checkout, never copied) against this
# library's OpenCorr-shaped headers: the reference's FFTCC2D lifecycle test,
# compiled unchanged, as a drop-in check.  Output only into oracle/_ref/.
#   make -f oracle/ref_probe.mk REF=/root/reference
REF ?= /root/reference
ROOT := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT := $(ROOT)/oracle/_ref/library_boundary_probe

$(OUT): $(REF)/smoke/library_boundary_probe.cpp $(ROOT)/include/fasticgn_opencorr.hpp $(ROOT)/include/fasticgn.h
	mkdir -p $(ROOT)/oracle/_ref
	g++ -std=c++17 -O1 -fopenmp -DOPENCORR_HAS_FFTW -I$(ROOT)/include/opencorr -I$(ROOT)/include \
	    $(REF)/smoke/library_boundary_probe.cpp -L$(ROOT)/fasticgn_amd -lfasticgn \
	    -Wl,-rpath,'$$ORIGIN/../../fasticgn_amd' -o $@

.PHONY: clean
clean:
	rm -f $(OUT)
