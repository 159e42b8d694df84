# ref.mk -- compiles the reference's dense CPU path from its own sources
# (read in place under REF, never copied) into oracle/_ref/, plus the golden
# harness oracle/ref_harness.cpp linked against it.  Test infrastructure only.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags follow the reference Makefile's release CPU build (Makefile:8, 45-47):
# g++ -std=c++11 -O3 -DNDEBUG, .cu compiled as C++.
REF ?= /root/reference
OUT := oracle/_ref
CXX ?= g++
CXXFLAGS := -std=c++11 -O3 -DNDEBUG -w -I$(REF)/include

OBJS := $(OUT)/suqa.o $(OUT)/system.o $(OUT)/Rand.o $(OUT)/io.o

all: $(OUT)/qms $(OUT)/evo $(OUT)/qsa $(OUT)/ref_harness $(OUT)/qms_gatecount

$(OUT):
	mkdir -p $(OUT)

$(OUT)/%.o: $(REF)/src/%.cu | $(OUT)
	$(CXX) $(CXXFLAGS) -x c++ -c $< -o $@

$(OUT)/%.o: $(REF)/src/%.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/qms: $(OUT)/qms.o $(OBJS)
	$(CXX) $^ -o $@

$(OUT)/evo: $(OUT)/evo.o $(OBJS)
	$(CXX) $^ -o $@

$(OUT)/qsa: $(OUT)/qsa.o $(OBJS)
	$(CXX) $^ -o $@

$(OUT)/ref_harness: oracle/ref_harness.cpp $(OBJS)
	$(CXX) $(CXXFLAGS) oracle/ref_harness.cpp $(OBJS) -o $@

# the same qms driver with the gate-count instrumentation compiled in (-DGATECOUNT)
GC_OBJS := $(OUT)/gc/qms.o $(OUT)/gc/suqa.o $(OUT)/gc/system.o $(OUT)/gc/Rand.o $(OUT)/gc/io.o

$(OUT)/gc:
	mkdir -p $(OUT)/gc

$(OUT)/gc/%.o: $(REF)/src/%.cu | $(OUT)/gc
	$(CXX) $(CXXFLAGS) -DGATECOUNT -x c++ -c $< -o $@

$(OUT)/gc/%.o: $(REF)/src/%.cpp | $(OUT)/gc
	$(CXX) $(CXXFLAGS) -DGATECOUNT -c $< -o $@

$(OUT)/qms_gatecount: $(GC_OBJS)
	$(CXX) $^ -o $@

.PHONY: all
