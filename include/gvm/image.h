// Image being optimised plus its per-image line-search functions: the
// reference's include/classes/image.cuh:4-31.  newP / evaluateXt point at
// the positivity (particular*) or plain (default*) variants
// (src/functions.cu:4566-4592, src/mfs.cu:794-811).
#pragma once

// newP and evaluateXt keep the reference's float step (image.cuh:4-7), so a
// mapping filled the reference's way (`m.newP = particularNewP;`, which picks
// the float overload below) works unchanged.  newPD and evaluateXtD are
// additions: the same functions with the step as a double, which is what
// Brent's minimisation returns here.  The optimisers call the double forms
// when they are set and the float forms (with the step rounded to float, as
// in the reference) otherwise; makeFunctionMapping sets both.
typedef struct functionMap {
  void (*newP)(float*, float*, float, int);
  void (*evaluateXt)(float*, float*, float*, float, int);
  void (*newPD)(float*, float*, double, int) = nullptr;
  void (*evaluateXtD)(float*, float*, float*, double, int) = nullptr;
} imageMap;

class Image {
 public:
  Image(float* image, int image_count) : image(image), image_count(image_count) {}
  int getImageCount() { return image_count; }
  float* getImage() { return image; }
  float* getErrorImage() { return error_image; }
  imageMap* getFunctionMapping() { return functionMapping; }
  void setImageCount(int i) { this->image_count = i; }
  void setErrorImage(float* f) { this->error_image = f; }
  void setImage(float* i) { this->image = i; }
  void setFunctionMapping(imageMap* f) { this->functionMapping = f; }

 private:
  float* image;
  int image_count;
  float* error_image = nullptr;
  imageMap* functionMapping = nullptr;
};

// src/functions.cu:4566-4592: positivity on image 0 (floor -eta*MINPIX),
// none on the others.  Float-step overloads (the reference's signatures) and
// double-step ones.
void defaultNewP(float* p, float* xi, float xmin, int image);
void particularNewP(float* p, float* xi, float xmin, int image);
void defaultEvaluateXt(float* xt, float* pcom, float* xicom, float x, int image);
void particularEvaluateXt(float* xt, float* pcom, float* xicom, float x, int image);
void defaultNewP(float* p, float* xi, double xmin, int image);
void particularNewP(float* p, float* xi, double xmin, int image);
void defaultEvaluateXt(float* xt, float* pcom, float* xicom, double x, int image);
void particularEvaluateXt(float* xt, float* pcom, float* xicom, double x, int image);
// Builds the mapping of src/mfs.cu:794-811 (owned by the caller, delete[]).
imageMap* makeFunctionMapping(int image_count, bool nopositivity);
