# Builds the reference's own CPU Siddon path (APIRL recon/ + data/ + utils/ + the
# vendored medcon Interfile reader) from its sources where they lie under
# /root/reference, plus a small driver of ours that exercises that library, into
# oracle/_ref/.  Nothing from /root/reference is copied into this repository:
# the compiler reads the sources in place and only the built objects land in
# oracle/_ref/ (git-ignored).
#
#   make -f oracle/ref.mk            # from the repository root
#
# Flags follow the reference's CMake Release build (CMakeLists.txt:23-33: OpenMP
# on; CMAKE_BUILD_TYPE Release -> -O3 -DNDEBUG).  No -march: the x86-64 baseline
# has no FMA, so float expressions round exactly as written.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
CC       ?= gcc
OPT      ?= -O3 -DNDEBUG
INC      := -I$(REF)/recon/inc -I$(REF)/data/inc -I$(REF)/utils/inc -I$(REF)/data/medcon
CXXFLAGS := $(OPT) -fopenmp -fPIC -w -D_CRT_SECURE_NO_WARNINGS $(INC)
CFLAGS   := $(OPT) -fopenmp -fPIC -w $(INC)

# recon/: the projector, Siddon and EM drivers the path needs (recon/recon.cmake
# globs every *.cpp; the other projectors are independent translation units and
# are left out because nothing on the path calls them).
RECON_SRC := Siddon.cpp SiddonProjector.cpp Projector.cpp Mlem.cpp Mlem2d.cpp \
             Mlem2dMultiple.cpp MlemSinogram3d.cpp OsemSinogram3d.cpp
DATA_SRC  := $(notdir $(wildcard $(REF)/data/src/*.cpp))
UTILS_SRC := $(notdir $(wildcard $(REF)/utils/src/*.cpp))
MEDCON_SRC:= $(notdir $(wildcard $(REF)/data/medcon/*.c))

OBJS := $(addprefix $(OUT)/recon_,$(RECON_SRC:.cpp=.o)) \
        $(addprefix $(OUT)/data_,$(DATA_SRC:.cpp=.o)) \
        $(addprefix $(OUT)/utils_,$(UTILS_SRC:.cpp=.o)) \
        $(addprefix $(OUT)/medcon_,$(MEDCON_SRC:.c=.o))

all: $(OUT)/libapirl_ref.so $(OUT)/ref_driver

$(OUT):
	mkdir -p $(OUT)

$(OUT)/recon_%.o: $(REF)/recon/src/%.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@
$(OUT)/data_%.o: $(REF)/data/src/%.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@
$(OUT)/utils_%.o: $(REF)/utils/src/%.cpp | $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@
$(OUT)/medcon_%.o: $(REF)/data/medcon/%.c | $(OUT)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/libapirl_ref.so: $(OBJS)
	$(CXX) -shared -fopenmp -o $@ $^ -lm

# -fno-access-control: the driver swaps Image's private pixel pointer for a guarded
# copy (see ref_driver.cpp); the reference's headers stay untouched.
$(OUT)/ref_driver: oracle/ref_driver.cpp $(OUT)/libapirl_ref.so
	$(CXX) $(OPT) -fopenmp -fno-access-control $(INC) -w -o $@ $< -L$(OUT) -lapirl_ref -Wl,-rpath,'$$ORIGIN' -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean

# The reference's own EM drivers with the MI355X projector plugged in through its
# reference-side binding (integration/HipSiddonProjector.cpp): the drop-in check.
HIPLIB := apirl_amd/libapirl_hip.so
hip: $(OUT)/ref_hip_driver

$(OUT)/ref_hip_driver: oracle/ref_driver.cpp integration/HipSiddonProjector.cpp integration/HipSiddonProjector.h $(OUT)/libapirl_ref.so $(HIPLIB)
	$(CXX) $(OPT) -fopenmp -fno-access-control -DWITH_HIP $(INC) -Iinclude -Iintegration -w -o $@ oracle/ref_driver.cpp integration/HipSiddonProjector.cpp -L$(OUT) -lapirl_ref -Lapirl_amd -lapirl_hip -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../apirl_amd'

.PHONY: hip
