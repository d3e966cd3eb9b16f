# Builds the reference's own CPU Poseidon + Merkle path (okx/zeknox native/)
# straight from its source files under /root/reference, into oracle/_ref/.
# It is only a checker for the oracle restatement; nothing here is shipped.
# The reference's Merkle dispatch (merkle/merkle.cpp) references every hasher,
# so all hasher translation units are compiled (as C++, they have host paths).
REF ?= /root/reference/native
OUT ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
SRCS = $(REF)/merkle/merkle.cpp $(REF)/poseidon/poseidon.cpp \
       $(REF)/keccak/keccak.cu $(REF)/poseidon2/poseidon2.cu \
       $(REF)/monolith/monolith.cu $(REF)/poseidon/poseidon_bn128.cu \
       $(REF)/poseidon/element_bn128.cu
CXXFLAGS = -x c++ -std=c++17 -O2 -fPIC -fopenmp -DFEATURE_GOLDILOCKS -DEXPOSE_C_INTERFACE -I$(REF) -w

all: $(OUT)/libzkref.so

$(OUT)/libzkref.so: $(SRCS)
	@mkdir -p $(OUT)
	g++ $(CXXFLAGS) -shared -Wl,--no-undefined $(SRCS) -o $@ -fopenmp

clean:
	rm -f $(OUT)/libzkref.so
.PHONY: all clean
