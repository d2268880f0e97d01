# Compiles Selas's own source files, where they lie under /root/reference, together with
# ref_harness.cpp into _ref/libselas_ref.so (test infrastructure; outputs only into _ref/).
#
# Flags: Selas has no Linux build target; -DIsOsx_=1 selects its POSIX platform macros
# (Source/Core/SystemLib/BasicTypes.h:19-26). MemoryAllocation.cpp calls posix_memalign without
# including <stdlib.h>, so that one file is compiled with -include stdlib.h (a system header).
# Nothing is stubbed: every symbol resolves from these files (linked with --no-undefined).
REF ?= /root/reference
CORE := $(REF)/Source/Core
APP := $(REF)/Source/Applications/Selas/Source
OUT := _ref
CXX ?= g++
CXXFLAGS := -std=c++14 -O2 -fPIC -ffp-contract=off -DIsOsx_=1 -w -ffunction-sections -fdata-sections \
            -I$(CORE) -I$(APP) -I$(REF)/Middleware -I$(REF)/Middleware/embree-latest/include \
            -I$(REF)/Middleware/ptex/Include

SRCS := $(CORE)/Shading/Disney.cpp $(CORE)/Shading/Ggx.cpp $(CORE)/Shading/Fresnel.cpp \
        $(CORE)/MathLib/FloatFuncs.cpp $(CORE)/MathLib/FloatStructs.cpp $(CORE)/MathLib/Trigonometric.cpp \
        $(CORE)/MathLib/Sampler.cpp $(CORE)/MathLib/Random.cpp $(CORE)/MathLib/CorrelatedMultiJitter.cpp \
        $(CORE)/MathLib/GeometryIntersection.cpp $(CORE)/GeometryLib/Camera.cpp \
        $(CORE)/GeometryLib/RectangulerLightSampler.cpp $(CORE)/GeometryLib/Disc.cpp \
        $(CORE)/GeometryLib/AxisAlignedBox.cpp $(CORE)/SystemLib/Memory.cpp $(CORE)/SystemLib/JsAssert.cpp \
        $(CORE)/IoLib/Serializer.cpp $(APP)/VCMHashGrid.cpp
# the resource files' serializers (SerializeToBinary / AttachToBinary over each resource's own
# Serialize function) and the name hash. ModelResource.cpp, SceneResource.cpp and
# SubsceneResource.cpp also hold the Embree scene set-up; the library is linked with
# --gc-sections and exports only the harness's ref_* entry points, so the functions that call
# Embree are dropped as unreferenced (nothing is stubbed: --no-undefined still holds).
SRCS += $(CORE)/IoLib/BinarySerializers.cpp $(CORE)/IoLib/SizeSerializer.cpp \
        $(CORE)/SceneLib/ModelResource.cpp $(CORE)/SceneLib/SceneResource.cpp \
        $(CORE)/SceneLib/SubsceneResource.cpp $(CORE)/SceneLib/ImageBasedLightResource.cpp \
        $(CORE)/UtilityLib/MurmurHash.cpp
# the vendored Ptex 2.x reader and filters (PtexFilter::getFilter, PtexTexture::open); zlib is the
# system's
PTEX := $(REF)/Middleware/ptex/Include
SRCS += $(PTEX)/PtexReader.cpp $(PTEX)/PtexCache.cpp $(PTEX)/PtexFilters.cpp $(PTEX)/PtexHalf.cpp \
        $(PTEX)/PtexSeparableFilter.cpp $(PTEX)/PtexSeparableKernel.cpp $(PTEX)/PtexTriangleFilter.cpp \
        $(PTEX)/PtexTriangleKernel.cpp $(PTEX)/PtexUtils.cpp
OBJS := $(patsubst $(REF)/%.cpp,$(OUT)/obj/%.o,$(SRCS)) $(OUT)/obj/MemoryAllocation.o $(OUT)/obj/ref_harness.o

all: $(OUT)/libselas_ref.so

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/MemoryAllocation.o: $(CORE)/SystemLib/MemoryAllocation.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -include stdlib.h -c $< -o $@

$(OUT)/obj/ref_harness.o: ref_harness.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libselas_ref.so: $(OBJS) ref_exports.map
	$(CXX) -shared -o $@ $(OBJS) -lz -Wl,--no-undefined -Wl,--gc-sections -Wl,--version-script=ref_exports.map

.PHONY: all
