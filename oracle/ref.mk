# Builds the reference's own C++ sources for the HMM path in place (never
# copied) into oracle/_ref/libwpref.so with the release flags (-O2 -flto,
# cc_code/makefile:121), with oracle/ref_shim.cc exposing a
# plain C entry set for the tests. levmar is compiled from its C sources. The
# prebuilt MR-MPI objects and the dedup code that needs them are not used.
REF_SRC = $(REF)/src
REF_DIRS = hmm/hmm hmm/precomputations hmm/state-vector hmm/step tensor \
           parameterization/fit parameterization/model parameterization/settings \
           common classifiers fitters
REF_CC = $(filter-out %.test.cc,$(foreach d,$(REF_DIRS),$(wildcard $(REF_SRC)/$(d)/*.cc))) \
         $(REF_SRC)/util/kd-range.cc $(REF_SRC)/util/levmar-helper.cc \
         $(REF_SRC)/util/range.cc $(REF_SRC)/util/time.cc \
         $(REF_SRC)/simulation/generate-dye-track.cc \
         $(REF_SRC)/simulation/generate-radiometry.cc \
         $(REF_SRC)/simulation/generate-radiometries.cc \
         $(REF_SRC)/io/radiometries-io.cc $(REF_SRC)/io/dye-tracks-io.cc \
         $(REF_SRC)/io/scored-classifications-io.cc
LEVMAR_DIR = $(REF)/extern/levmar-2.6
LEVMAR_C = $(addprefix $(LEVMAR_DIR)/,lm.c Axb.c misc.c lmlec.c lmbc.c lmblec.c lmbleic.c)
REF_OUT = $(HERE)_ref
REF_INC = -iquote$(REF_SRC) -I$(REF)/extern/nlohmann-json-3.11.2/ -I$(LEVMAR_DIR)

ref: $(REF_OUT)/libwpref.so

$(REF_OUT)/liblevmar_pic.a: $(LEVMAR_C)
	mkdir -p $(REF_OUT)/levmar_obj
	cd $(REF_OUT)/levmar_obj && $(CC) -O2 -fPIC -c $(LEVMAR_C) -I$(LEVMAR_DIR)
	ar crs $@ $(REF_OUT)/levmar_obj/*.o

$(REF_OUT)/libwpref.so: $(REF_CC) $(HERE)ref_shim.cc $(REF_OUT)/liblevmar_pic.a
	$(CXX) -std=gnu++17 -O2 -flto -fPIC -fopenmp -shared $(REF_INC) -I$(HERE)../include \
	    -o $@ $(REF_CC) $(HERE)ref_shim.cc $(REF_OUT)/liblevmar_pic.a -lm
